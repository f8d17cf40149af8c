// pa_chain.hip -- CDNA4 (gfx950) chain kernel for PhyloAcc-ST's per-element MCMC.
//
// The fixed-(Z, r) log-likelihood kernel is in pa_engine.hip.
//
// One workgroup owns one element at a time and runs its three model chains
// (M0, M1, M2 = resZ 0, 2, 1; src/PhyloAcc-ST/main.cpp:466-494) start to
// finish; workgroups pull elements from a device work queue (persistent
// kernel).  Within a workgroup:
//   * threads stride over site patterns (pruning) or sites (ancestral
//     sampling); every thread keeps the same patterns through a whole
//     post-order pass, so no barrier is needed between tree nodes;
//   * conditional likelihoods are pattern-major in HBM scratch, two buffers
//     per internal node (current / proposal) so an accepted Metropolis move
//     flips a per-node bit instead of copying partials;
//   * partials are linear-space with a power-of-two rescale per node; the
//     cumulative exponent rides along as an int; GAPT marks all-gap subtrees
//     (the reference's Tg >= num_base, src/PhyloAcc-ST/bpp_c.hpp:417-428);
//   * a cherry (a non-root internal node whose children are both leaves) is
//     never stored: its parent and the ancestral sampler recompute its
//     scaled partial from the two leaf codes with the same operations, so
//     results are unchanged and a pass saves the cherry's store, its reload
//     and one dependent memory round trip;
//   * per-branch 4x4 transition matrices of the current rate class live in
//     LDS (read as broadcasts); FP64 is used for P(t) builds, log tables,
//     emissions, the Z forward-backward, likelihood sums and MH ratios.
// The chain context and the model constants live in LDS behind
// address-space-3 pointers, and every sweep phase is a separate non-inlined
// function, so each phase gets its own register allocation instead of one
// allocation spanning the whole persistent loop.
// No MFMA: a K=4 matvec per pattern is not a dense contraction.
#include <hip/hip_runtime.h>
#include <type_traits>
#include <algorithm>
#include <stdint.h>
#include "pa_common.h"
#include "pa_rng.h"
#include "pa_chain.h"
#include "pa_device.h"

namespace pa {
namespace chain {

template <typename R>
struct Ctx {
    // element
    int S, N, I, GG, P, Ppad, root, n_nodes, n_int, n_up, n_down, has_amb;
    const GBL uint8_t *codes;
    const GBL int32_t *wts, *site;
    // workgroup scratch in HBM
    GBL V4<R> *parts;      // [2][I][Ppad]
    GBL int32_t *exps;     // [2][I][Ppad]
    GBL double *plin;      // [3][N][16]  P_{j->i} at [j*4+i]
    GBL double *logt;      // [3][N][16]  log P_{j->i} at [i*4+j]
    GBL uint32_t *amb;     // [S][24]
    const GBL double *dist2;   // [N] branch lengths (element node table)
    // LDS arrays
    LDS int16_t *par, *nodes, *inodes, *up, *upoff, *down, *downoff, *tz;
    LDS uint32_t *desc;    // [I][4] static pruning descriptor of each internal node (node_desc)
    LDS int32_t *plist;    // [I] nodes a pass visits, level order: s | level << 16
    LDS uint8_t *flags, *fixZ, *dirty, *cur, *changed;
    LDS uint8_t *pb;       // [N] per-pass buffer / matrix bits (pass_bits); shares storage with changed
    LDS int8_t *Z, *maxZ;
    LDS double *E, *beta;
    LDS R *PTc, *PTp;      // [N][16], PT[i*4+j] = P_{j->i}
    LDS uint32_t *hist, *cnt9;
    LDS uint8_t *tgs;
    // chain scalars
    double loglik, nrate, crate, g, l, l2, prop_n, prop_c, proposal, r, max_loglik;
    double TMf[9], TMn[9];  // log_TM_Int for fixZ / free nodes, [i*3+j] = log P(child i | parent j)
    double red[16];
    unsigned long long bytes;   // algorithmic HBM bytes of this element's passes (see prune / sample_sites)
    int accept_n, accept_c, max_m, skip, accept, fail, work, nch, n_vis;   // fail: PA_F_* bits
};

template <typename R>
using CP = LDS Ctx<R> *;
using MP = LDS const ModelDev *;

__host__ __device__ inline size_t al(size_t x, size_t a) { return (x + a - 1) & ~(a - 1); }

// Dynamic-LDS layout: returns the byte size; BIND also points the context's
// array fields into the block.  The context may sit at LDS address 0, which
// compares equal to a null pointer, so binding is selected at compile time and
// never by testing the context pointer.
template <typename R, bool BIND>
__host__ __device__ size_t smem_arrays(int N, int S, int BS, LDS char *base, CP<R> c)
{
    const int I = N - S;
    size_t o = 0;
#define PA_TAKE(field, T, count)                                                   \
    do {                                                                           \
        o = al(o, sizeof(T) < 4 ? 4 : sizeof(T));                                  \
        if constexpr (BIND) c->field = (LDS T *)(base + o);                        \
        o += sizeof(T) * (size_t)(count);                                          \
    } while (0)
    // matrices first: rows of PT are read as 4-vectors (32-byte aligned base)
    PA_TAKE(PTc, R, 16 * N);
    // Phases that never overlap share storage:
    //   A: per-thread ancestral states (sample_sites) | emissions and Z
    //      messages (emissions, update_Z) | Z-transition counts
    //      (sample_transition);
    //   B: per-branch transition counters (sample_sites, emissions) |
    //      proposal matrices (rate updates).  sample_sites clears the
    //      counters before it counts.
    {
        const size_t szA = std::max({(size_t)(I > 0 ? I : 1) * BS, (size_t)6 * N * sizeof(double), (size_t)20 * 4});
        const size_t szB = std::max((size_t)16 * N * sizeof(uint32_t), (size_t)16 * N * sizeof(R));
        o = al(o, 32);
        if constexpr (BIND) {
            c->PTp = (LDS R *)(base + o);
            c->hist = (LDS uint32_t *)(base + o);
        }
        o += al(szB, 32);
        if constexpr (BIND) {
            c->E = (LDS double *)(base + o);
            c->beta = (LDS double *)(base + o) + 3 * N;
            c->tgs = (LDS uint8_t *)(base + o);
            c->cnt9 = (LDS uint32_t *)(base + o);
        }
        o += szA;
    }
    o = al(o, 16);      // descriptors are read as 16-byte vectors
    if constexpr (BIND) c->desc = (LDS uint32_t *)(base + o);
    o += (size_t)16 * (I > 0 ? I : 1);
    PA_TAKE(plist, int32_t, I > 0 ? I : 1);
    PA_TAKE(par, int16_t, N);
    PA_TAKE(nodes, int16_t, N);
    PA_TAKE(inodes, int16_t, N);
    PA_TAKE(up, int16_t, N);
    PA_TAKE(upoff, int16_t, N + 1);
    PA_TAKE(down, int16_t, N);
    PA_TAKE(downoff, int16_t, N + 1);
    PA_TAKE(tz, int16_t, N);
    PA_TAKE(flags, uint8_t, N);
    PA_TAKE(fixZ, uint8_t, N);
    PA_TAKE(dirty, uint8_t, N);
    PA_TAKE(cur, uint8_t, N);
    PA_TAKE(changed, uint8_t, N);
    if constexpr (BIND) c->pb = c->changed;   // changed lives inside update_Z, pb inside prune
    PA_TAKE(Z, int8_t, N);
    PA_TAKE(maxZ, int8_t, N);
#undef PA_TAKE
    return al(o, 16);
}

size_t chain_dyn_smem_bytes(int N, int S, int BS, int fp64)
{
    return fp64 ? smem_arrays<double, false>(N, S, BS, nullptr, nullptr) : smem_arrays<float, false>(N, S, BS, nullptr, nullptr);
}

size_t chain_smem_bytes(int N, int S, int BS, int fp64)
{
    return (fp64 ? smem_arrays<double, false>(N, S, BS, nullptr, nullptr)
                 : smem_arrays<float, false>(N, S, BS, nullptr, nullptr)) +
           (fp64 ? sizeof(Ctx<double>) : sizeof(Ctx<float>)) + sizeof(ModelDev);
}

// Byte offset of the exponents behind the partials of an element of Ppad
// padded patterns ([2][I][Ppad] 4-vectors of R).
__host__ __device__ inline size_t chain_exps_off(int N, int S, int Ppad, int rs)
{
    return al((size_t)2 * (N - S) * Ppad * 4 * rs, 256);
}

size_t chain_part_bytes(int N, int S, int Ppad, int fp64)
{
    return chain_exps_off(N, S, Ppad, fp64 ? 8 : 4) + al((size_t)2 * (N - S) * Ppad * 4, 256);
}

// Per-workgroup HBM scratch layout (bytes), shared with the host: the rate
// class tables and ambiguity counters, then the partial region.
size_t slot_scratch_bytes(int N, int S, int Pcap, int fp64, size_t *off)
{
    size_t o = 0;
    off[0] = o; o = al(o + (size_t)3 * N * 16 * 8, 256);          // plin
    off[1] = o; o = al(o + (size_t)3 * N * 16 * 8, 256);          // logt
    off[2] = o; o = al(o + (size_t)S * 24 * 4, 256);              // amb
    off[3] = o; o += chain_part_bytes(N, S, Pcap, fp64);          // parts, exps
    return o;
}

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ void lds_add(LDS uint32_t *p, uint32_t v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// P(t) = V diag(exp(lam t)) Vi, clamped at 0; out[j*4+i] = P_{j->i}
template <typename OutPtr>
__device__ __forceinline__ void ptrans(MP M, double t, OutPtr out)
{
    double e[4];
#pragma unroll
    for (int k = 0; k < 4; k++) e[k] = exp(M->lam[k] * t);
#pragma unroll 1
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) s += M->V[j * 4 + k] * e[k] * M->Vi[k * 4 + i];
            out[j * 4 + i] = s < 0 ? 0.0 : s;
        }
}

// class tables for branch s, class k (0 null, 1 conserved, 2 accelerated) at rate r.
// mode 0: model start (dist2 > 1e-10 -> P(dist2*r), else identity; bpp_c.cpp:187-228)
// mode 1: refresh after a rate change (dist2 < 1e-10 keeps the old table; bpp_c.cpp:1115-1131)
template <typename R>
__device__ __noinline__ void class_table(MP M, CP<R> c, int k, int s, double r, int mode)
{
    const double d = c->dist2[s];
    GBL double *P = c->plin + ((size_t)k * c->N + s) * 16;
    GBL double *L = c->logt + ((size_t)k * c->N + s) * 16;
    if (mode == 1 && d < 1e-10) return;
    if (d > 1e-10 || mode == 1) ptrans(M, d * r, P);
    else
        for (int i = 0; i < 16; i++) P[i] = (i % 5 == 0) ? 1.0 : 0.0;
#pragma unroll 1
    for (int q = 0; q < 16; q++) L[q] = log(P[(q & 3) * 4 + (q >> 2)]);
}

// PT of the current class of every branch: all branches (only_cls = -1),
// those with Z == only_cls, or (REFRESH_CHANGED) those whose Z the last Z
// update changed
constexpr int REFRESH_CHANGED = -2;
template <typename R, int BS, int SH>
__device__ __noinline__ void refresh_PTc(CP<R> c, int only_cls)
{
    for (int q = threadIdx.x; q < c->n_nodes - 1; q += BS) {
        const int s = c->nodes[q];
        const int z = c->Z[s];
        if (only_cls >= 0 && z != only_cls) continue;
        if (only_cls == REFRESH_CHANGED && !c->changed[s]) continue;
        const int k = z == 2 ? 2 : (z == 0 ? 0 : 1);
        const GBL double *P = c->plin + ((size_t)k * c->N + s) * 16;
        LDS R *dst = c->PTc + 16 * s;
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) dst[i * 4 + j] = (R)P[j * 4 + i];
    }
}

// ----------------------------------------------------------------- pruning

enum { PR_FULL = 0, PR_DIRTY = 1, PR_PROP = 2 };

// Where one child of a node comes from in a pass: a leaf's code row, an
// internal child's partial / exponent rows (offsets into the slot arrays), or
// a cherry's two leaf code rows; and the branch matrices that apply.  Every
// field is the same on all lanes (scalar registers).
enum { KD_LEAF = 0, KD_STORED = 1, KD_CHERRY = 2, KD_KIND = 3, KD_MISS = 4, KD_MISS_A = 8, KD_MISS_B = 16 };

template <typename R>
struct KidDesc {
    int bits;             // kind (KD_KIND), the child missing, a cherry's leaves missing
    int off, offB;        // code row / partial offset; cherry: code rows of its two leaves
    const LDS R *src;     // the child's branch matrix
    const LDS R *sA, *sB; // cherry: its leaves' branch matrices
};

template <typename R>
struct NodeDesc {
    int dst;              // 2 * (offset of the node's partial / exponent rows in the written buffer) + missing
    KidDesc<R> k[2];
};

// Static pruning descriptor of an internal node, built when the element is
// bound (node_desc): word 0 = its children x0 | x1 << 16; words 1 and 2 = the
// two leaves of x0 (resp. x1) when that child is a cherry, yA | yB << 16;
// word 3 = DB_* bits: each child's kind (KD_LEAF / KD_STORED / KD_CHERRY),
// which children and cherry leaves are missing, and whether the node is.
enum { DB_KIND1_SHIFT = 2, DB_MISS0 = 1 << 4, DB_MISS1 = 1 << 5, DB_MISS_A0 = 1 << 6, DB_MISS_B0 = 1 << 7,
       DB_MISS_A1 = 1 << 8, DB_MISS_B1 = 1 << 9, DB_MISS = 1 << 10 };

// Per-pass bits of node y (pass_bits): PB_BUF = the buffer holding y's
// partial in this pass (read as a child, written as a visited node);
// PB_PROP = y's branch takes the proposal matrix.
enum { PB_BUF = 1, PB_PROP = 2 };

template <typename R>
__device__ __forceinline__ KidDesc<R> kid_from(int kind, int miss, int missA, int missB, int x, int px, int yA, int pA,
                                               int yB, int pB, const LDS R *PTc, const LDS R *PTp, int S, int Ppad,
                                               int bufI)
{
    KidDesc<R> d;
    d.src = ((px & PB_PROP) ? PTp : PTc) + 16 * x;
    d.offB = 0;
    d.sA = d.sB = PTc;
    if (kind == KD_LEAF) {
        d.bits = KD_LEAF | (miss ? KD_MISS : 0);
        d.off = x * Ppad;
    } else if (kind == KD_CHERRY) {
        d.bits = KD_CHERRY | (miss ? KD_MISS : 0) | (missA ? KD_MISS_A : 0) | (missB ? KD_MISS_B : 0);
        d.off = yA * Ppad;
        d.offB = yB * Ppad;
        d.sA = ((pA & PB_PROP) ? PTp : PTc) + 16 * yA;
        d.sB = ((pB & PB_PROP) ? PTp : PTc) + 16 * yB;
    } else {
        d.bits = KD_STORED | (miss ? KD_MISS : 0);
        d.off = (px & PB_BUF) * bufI + (x - S) * Ppad;
    }
    return d;
}

template <typename R>
__device__ __forceinline__ ChildIn<R> kid_load(const KidDesc<R> &d, const GBL uint8_t *codes, const GBL V4<R> *parts,
                                               const GBL int32_t *exps, int p)
{
    // a missing child contributes nothing, but its gap state still decides
    // whether the parent is all-gap at this pattern, so its code / exponent
    // is read either way; a cherry's two codes are packed into x
    ChildIn<R> in;
    in.v.x = in.v.y = in.v.z = in.v.w = (R)0;
    const int kind = d.bits & KD_KIND;
    if (kind == KD_LEAF) {
        in.x = codes[d.off + p];
    } else if (kind == KD_CHERRY) {
        in.x = codes[d.off + p] | (codes[d.offB + p] << 8);
    } else {
        in.x = exps[d.off + p];
        if (!(d.bits & KD_MISS)) in.v = ld4<R>(parts + d.off + p);
    }
    return in;
}

// Pattern-loop shapes the autotuner picks among.  U: visited nodes pruned
// together per round (all their loads are issued before the first is used);
// K: internal nodes whose partials one sampling chunk loads before drawing.
// Shape 0 is the default: two nodes in FP32 and one in FP64 fit the
// 128-register budget of the chain kernel without spilling in the pattern
// loop (groups of 3 spill and ran 3 % slower on the 50,000-element
// 43-species workload).
template <typename R, int SH>
struct Shape;
template <> struct Shape<float, 0> { static constexpr int U = 2, K = 4; };
template <> struct Shape<float, 1> { static constexpr int U = 1, K = 4; };
template <> struct Shape<float, 2> { static constexpr int U = 2, K = 8; };
template <> struct Shape<double, 0> { static constexpr int U = 1, K = 2; };
template <> struct Shape<double, 1> { static constexpr int U = 2, K = 2; };
template <> struct Shape<double, 2> { static constexpr int U = 1, K = 4; };

// One pruning pass (BPP_C::Update_Tg(tosample=false), src/PhyloAcc-ST/bpp_c.cpp:644-671,
// and log_f_Xz, bpp_c.cpp:860-928).  FULL: every internal node of the tree
// (missing nodes only record their GAPT flag) into the current buffer.
// DIRTY: nodes with dirty[] set, in place.  PROP: dirty nodes into the other
// buffer, with PTp for branches whose Z is cls.  Cherries are skipped: their
// parent recomputes them (cherry_in) with the matrices the pass would have
// used.  Returns the element's log-likelihood (FP64), identical on every
// thread.
//
// Nodes are taken by tree level (the element's height-ordered list), so the
// nodes of one round never depend on each other: a round takes up to U
// visited nodes of one level and, for each pattern block, issues the loads
// of all their children before computing any of them.  A pass then costs
// about one memory round trip per round instead of one per node.  A thread
// owns the patterns p = tid (mod BS) of every node, so it only reads partials
// it wrote itself and the pass needs no barrier between rounds.  The pass
// first lists the nodes it visits (one ballot per 64 nodes of a level) and
// every node's buffer and matrix bits, so a round's scalar set-up is a few
// independent LDS reads: its list entries, its nodes' static descriptors
// (node_desc) and their children's bits.
template <typename R, int BS, int SH>
__device__ __noinline__ double prune(MP M, CP<R> c, int mode, int cls)
{
    const int tid = threadIdx.x;
    const int P = uni(c->P), Ppad = uni(c->Ppad), S = uni(c->S), n_up = uni(c->n_up);
    const int bufI = uni(c->I) * Ppad;
    GBL V4<R> *parts = uni_ptr(c->parts);
    GBL int32_t *exps = uni_ptr(c->exps);
    const GBL uint8_t *codes = uni_ptr(c->codes);
    const LDS R *PTc = uni_ptr(c->PTc), *PTp = uni_ptr(c->PTp);
    constexpr int U = Shape<R, SH>::U;
    unsigned long long bytes = 0;
    const LDS uint8_t *flags = uni_ptr(c->flags), *dirty = uni_ptr(c->dirty), *cur = uni_ptr(c->cur);
    const LDS int32_t *plist = uni_ptr(c->plist);
    // pass set-up: every node's buffer and matrix bits, and the list of
    // nodes the pass visits in level order (one wave compacts it by ballot)
    {
        const bool prop = mode == PR_PROP;
        const LDS int16_t *nodes = uni_ptr(c->nodes), *up = uni_ptr(c->up), *upoff = uni_ptr(c->upoff);
        const LDS int8_t *Zs = uni_ptr(c->Z);
        LDS uint8_t *pbw = uni_ptr(c->pb);
        LDS int32_t *plw = uni_ptr(c->plist);
        const int n_nodes = uni(c->n_nodes);
        for (int q = tid; q < n_nodes; q += BS) {
            const int y = nodes[q];
            const int cu = cur[y];
            pbw[y] = (uint8_t)(((prop && dirty[y]) ? 1 - cu : cu) | ((prop && Zs[y] == cls) ? PB_PROP : 0));
        }
        if (tid < 64) {
            int base = 0;
            for (int h = 0; h < n_up; h++) {
                const int qa = h == 0 ? 0 : upoff[h - 1], qb = upoff[h];
                for (int q0 = qa; q0 < qb; q0 += 64) {
                    const int q = q0 + tid;
                    bool v = false;
                    int y = 0;
                    if (q < qb) {
                        y = up[q];
                        const int f = flags[y];
                        v = !(f & NF_CHERRY) && (mode == PR_FULL || (dirty[y] && !(f & NF_MISSING)));
                    }
                    const unsigned long long bal = __ballot(v);
                    if (v) plw[base + __popcll(bal & ((1ull << tid) - 1))] = y | (h << 16);
                    base += __popcll(bal);
                }
            }
            if (tid == 0) c->n_vis = base;
        }
        __syncthreads();
    }
    const int n_vis = uni(c->n_vis);
    const LDS uint8_t *pb = uni_ptr(c->pb);
    const LDS uint32_t *desc = uni_ptr(c->desc);
    for (int k = 0; k < n_vis;) {
        {
            // up to U consecutive visited nodes of one level
            int ent[U], sel[U];
#pragma unroll
            for (int u = 0; u < U; u++) ent[u] = (k + u < n_vis) ? uni(plist[k + u]) : -1;
            int used = 1;
            sel[0] = ent[0] & 0xffff;
#pragma unroll
            for (int u = 1; u < U; u++) {
                sel[u] = -1;
                if (used == u && ent[u] >= 0 && (ent[u] >> 16) == (ent[0] >> 16)) { sel[u] = ent[u] & 0xffff; used++; }
            }
            k += used;
            NodeDesc<R> nd[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (sel[u] < 0) continue;
                const int s = sel[u];
                const LDS uint32_t *dd = desc + 4 * (s - S);
                const uint32_t w0 = uni(dd[0]), w1 = uni(dd[1]), w2 = uni(dd[2]), w3 = uni(dd[3]);
                const int x0 = w0 & 0xffff, x1 = w0 >> 16;
                const int yA0 = w1 & 0xffff, yB0 = w1 >> 16, yA1 = w2 & 0xffff, yB1 = w2 >> 16;
                const int ps = uni(pb[s]), p0 = uni(pb[x0]), p1 = uni(pb[x1]);
                const int pA0 = uni(pb[yA0]), pB0 = uni(pb[yB0]), pA1 = uni(pb[yA1]), pB1 = uni(pb[yB1]);
                nd[u].dst = 2 * ((ps & PB_BUF) * bufI + (s - S) * Ppad) + ((w3 & DB_MISS) ? 1 : 0);
                nd[u].k[0] = kid_from<R>(w3 & 3, w3 & DB_MISS0, w3 & DB_MISS_A0, w3 & DB_MISS_B0, x0, p0, yA0, pA0, yB0,
                                         pB0, PTc, PTp, S, Ppad, bufI);
                nd[u].k[1] = kid_from<R>((w3 >> DB_KIND1_SHIFT) & 3, w3 & DB_MISS1, w3 & DB_MISS_A1, w3 & DB_MISS_B1, x1,
                                         p1, yA1, pA1, yB1, pB1, PTc, PTp, S, Ppad, bufI);
                // write partial + exponent; read each child's partial + exponent, its leaf code, or a cherry's two codes
                bytes += (unsigned long long)P * (4 * sizeof(R) + 4);
#pragma unroll
                for (int j = 0; j < 2; j++)
                    bytes += (unsigned long long)P * ((nd[u].k[j].bits & KD_KIND) == KD_LEAF ? 1 : (nd[u].k[j].bits & KD_KIND) == KD_CHERRY ? 2 : 4 * sizeof(R) + 4);
            }
            for (int p = tid; p < P; p += BS) {
                ChildIn<R> in[U][2];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (sel[u] < 0) continue;
                    in[u][0] = kid_load<R>(nd[u].k[0], codes, parts, exps, p);
                    in[u][1] = kid_load<R>(nd[u].k[1], codes, parts, exps, p);
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    // one node at a time: keeps the scheduler from hoisting
                    // every node's matrices into registers at once
                    __builtin_amdgcn_sched_barrier(0);
                    if (sel[u] < 0) continue;
                    R a0 = 1, a1 = 1, a2 = 1, a3 = 1;
                    int e = 0;
                    bool g[2];
#pragma unroll
                    for (int j = 0; j < 2; j++) {
                        const KidDesc<R> &d = nd[u].k[j];
                        const int kind = d.bits & KD_KIND;
                        const bool miss = d.bits & KD_MISS;
                        ChildIn<R> cin = in[u][j];
                        if (kind == KD_CHERRY)
                            cin = cherry_in<R>(d.sA, d.sB, d.bits & KD_MISS_A, d.bits & KD_MISS_B, miss, cin.x & 255, cin.x >> 8);
                        child_apply<R>(kind == KD_LEAF, miss, d.src, cin, a0, a1, a2, a3, e, g[j]);
                    }
                    const int dst = nd[u].dst >> 1;
                    GBL int32_t *dste = exps + dst;
                    if (g[0] && g[1]) {
                        dste[p] = GAPT;
                    } else if (nd[u].dst & 1) {
                        dste[p] = 0;
                    } else {
                        const R mx = fmax(fmax(a0, a1), fmax(a2, a3));
                        int kx = 0;
                        frexp(mx, &kx);
                        a0 = ldexp(a0, -kx); a1 = ldexp(a1, -kx); a2 = ldexp(a2, -kx); a3 = ldexp(a3, -kx);
                        st4<R>(parts + dst + p, a0, a1, a2, a3);
                        dste[p] = e + kx;
                    }
                }
            }
        }
    }
    if (tid == 0) c->bytes += bytes;
    // root: log sum_j pi_j L_j + exponent * ln 2; all-gap columns contribute 0
    const int r = uni(c->root);
    double acc = 0;
    if (r >= S) {
        const int rb = uni(pb[r]) & PB_BUF;
        const GBL V4<R> *rp = parts + rb * bufI + (size_t)(r - S) * Ppad;
        const GBL int32_t *re = exps + rb * bufI + (size_t)(r - S) * Ppad;
        const GBL int32_t *wts = uni_ptr(c->wts);
        const double p0 = M->pi[0], p1 = M->pi[1], p2 = M->pi[2], p3 = M->pi[3];
        if (tid == 0) c->bytes += (unsigned long long)P * (4 * sizeof(R) + 4 + 4);   // root partial, exponent, weight
        for (int p = tid; p < P; p += BS) {
            const int ex = re[p];
            if (ex == GAPT) continue;
            const V4<R> a = ld4<R>(rp + p);
            const double L = p0 * (double)a.x + p1 * (double)a.y + p2 * (double)a.z + p3 * (double)a.w;
            acc += (double)wts[p] * (log(L) + (double)ex * 0.69314718055994530942);
        }
    }
    return block_sum(acc, c->red);
}

// ------------------------------------------------------ ancestral sampling

// BPP_C::Update_Tg(tosample=true) (src/PhyloAcc-ST/bpp_c.cpp:672-733) for every
// site, plus the per-branch transition counts that getEmission
// (bpp_c2.cpp:355-390) sums over.  Uniform for site g, sampling position t
// (0 = root, then internal nodes in reverse post-order): ANC stream, c1 = g.
template <typename R, int BS, int SH>
__device__ __noinline__ void sample_sites(MP M, CP<R> c, uint32_t c2, uint32_t c3)
{
    const int tid = threadIdx.x, S = uni(c->S), n_int = uni(c->n_int), GG = uni(c->GG), Ppad = uni(c->Ppad);
    const size_t bufI = (size_t)uni(c->I) * Ppad;
    const int r = uni(c->root);
    const uint32_t k0 = M->key0, k1 = M->key1;
    const double p0 = M->pi[0], p1 = M->pi[1], p2 = M->pi[2], p3 = M->pi[3];
    // the context's array pointers, read once: stores and atomics into LDS
    // arrays would otherwise make every use reload them
    const GBL V4<R> *parts = uni_ptr(c->parts);
    const GBL int32_t *exps = uni_ptr(c->exps);
    const GBL uint8_t *codes = uni_ptr(c->codes);
    const GBL int32_t *site = uni_ptr(c->site);
    GBL uint32_t *amb = uni_ptr(c->amb);
    LDS uint8_t *tgs = uni_ptr(c->tgs);
    LDS uint32_t *hist = uni_ptr(c->hist);
    const LDS int16_t *inodes = uni_ptr(c->inodes), *par = uni_ptr(c->par);
    const LDS uint8_t *flags = uni_ptr(c->flags), *cur = uni_ptr(c->cur);
    const LDS uint32_t *desc = uni_ptr(c->desc);
    const LDS R *PTc = uni_ptr(c->PTc);
    for (int i = tid; i < 16 * c->N; i += BS) hist[i] = 0;   // shares storage with the proposal matrices
    __syncthreads();
    if (tid == 0) {
        // per site: pattern index, root partial, each non-missing internal node's partial
        // (a cherry's two leaf codes), each leaf code
        int nint = 0, nch = 0, nleaf = 0;
        for (int k = 0; k < n_int; k++) {
            const uint8_t f = flags[inodes[k]];
            if (f & NF_MISSING) continue;
            if (f & NF_CHERRY) nch++; else nint++;
        }
        for (int q = 0; q < S; q++) nleaf += (flags[q] & NF_IN_TREE) && !(flags[q] & NF_MISSING);
        c->bytes += (unsigned long long)GG * (4 + (4 * sizeof(R) + 4) * nint + 2 * nch + nleaf);
    }
    for (int g = tid; g < GG; g += BS) {
        const int p = site[g];
        U4 blk = philox(0, (uint32_t)g, c2, c3, k0, k1);
        {
            const size_t off = cur[r] * bufI + (size_t)(r - S) * Ppad + p;
            double pr[4];
            if (exps[off] == GAPT) {
                pr[0] = p0; pr[1] = p1; pr[2] = p2; pr[3] = p3;
            } else {
                const V4<R> a = ld4<R>(parts + off);
                pr[0] = p0 * (double)a.x; pr[1] = p1 * (double)a.y; pr[2] = p2 * (double)a.z; pr[3] = p3 * (double)a.w;
            }
            tgs[(r - S) * BS + tid] = (uint8_t)categorical<4>(pr, u01(blk.x));
        }
        // internal nodes in reverse post-order, K at a time: a chunk's partial
        // loads do not depend on its draws, so they are all issued first.  A
        // cherry's slot holds its two leaf codes (low and second byte) until
        // its partial is recomputed.
        constexpr int K = Shape<R, SH>::K;
        int t = 1;
        for (int kb = n_int - 2; kb >= 0; kb -= K) {
            // node ids, flags, buffers and cherry leaves are the same on every
            // lane: they live in scalar registers
            int ex[K], sn[K], fl[K], kid[K];
            V4<R> av[K];
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int k = kb - j;
                ex[j] = GAPT;
                sn[j] = 0;
                fl[j] = NF_MISSING;
                kid[j] = 0;
                av[j].x = av[j].y = av[j].z = av[j].w = (R)0;
                if (k >= 0) {
                    const int s = uni(inodes[k]);
                    const int f = uni(flags[s]);
                    sn[j] = s;
                    fl[j] = f;
                    if (f & NF_CHERRY) {
                        const int kids = uni((int)desc[4 * (s - S)]);
                        kid[j] = kids;
                        if (!(f & NF_MISSING))
                            ex[j] = codes[(size_t)(kids & 0xffff) * Ppad + p] | (codes[(size_t)((uint32_t)kids >> 16) * Ppad + p] << 8);
                    } else if (!(f & NF_MISSING)) {
                        const size_t off = uni(cur[s]) * bufI + (size_t)(s - S) * Ppad + p;
                        ex[j] = exps[off];
                        av[j] = ld4<R>(parts + off);
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < K; j++, t++) {
                if (kb - j < 0) break;
                if ((t & 3) == 0) blk = philox((uint32_t)(t >> 2), (uint32_t)g, c2, c3, k0, k1);
                const int s = sn[j];
                if (fl[j] & NF_MISSING) continue;
                if ((fl[j] & NF_CHERRY) && ex[j] != GAPT) {
                    const int yA = kid[j] & 0xffff, yB = (int)((uint32_t)kid[j] >> 16);
                    const ChildIn<R> in = cherry_in<R>(PTc + 16 * yA, PTc + 16 * yB, uni(flags[yA]) & NF_MISSING,
                                                       uni(flags[yB]) & NF_MISSING, false, ex[j] & 255, ex[j] >> 8);
                    ex[j] = in.x;
                    av[j] = in.v;
                }
                if (ex[j] == GAPT) continue;     // missing node, or an all-gap subtree at this site
                const int tp = tgs[(uni(par[s]) - S) * BS + tid];
                const V4<R> a = av[j];
                const LDS R *PT = PTc + 16 * s;
                double pr[4] = {(double)PT[0 * 4 + tp] * (double)a.x, (double)PT[1 * 4 + tp] * (double)a.y,
                                (double)PT[2 * 4 + tp] * (double)a.z, (double)PT[3 * 4 + tp] * (double)a.w};
                const int ts = categorical<4>(pr, u01(pick(blk, t & 3)));
                tgs[(s - S) * BS + tid] = (uint8_t)ts;
                lds_add(&hist[s * 16 + tp * 4 + ts], 1u);
            }
        }
        // leaves, 8 codes loaded at a time
        for (int sb = 0; sb < S; sb += 8) {
            int cd[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int s = sb + j;
                cd[j] = PA_GAP_CODE;
                if (s < S) {
                    const int f = uni(flags[s]);
                    if ((f & NF_IN_TREE) && !(f & NF_MISSING)) cd[j] = codes[(size_t)s * Ppad + p];
                }
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int s = sb + j, code = cd[j];
                if (code == PA_GAP_CODE || code == PA_OTHER_CODE) continue;
                const int tp = tgs[(uni(par[s]) - S) * BS + tid];
                if (code < 4) lds_add(&hist[s * 16 + tp * 4 + code], 1u);
                else __hip_atomic_fetch_add(&amb[s * 24 + tp * 6 + (code - 6)], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    }
}

// getEmission (src/PhyloAcc-ST/bpp_c2.cpp:355-390): E[s][k] from the counts.
template <typename R, int BS, int SH>
__device__ __noinline__ void emissions(CP<R> c)
{
    const int N = c->N;
    for (int q = threadIdx.x; q < c->n_nodes - 1; q += BS) {
        const int s = c->nodes[q];
        if (c->flags[s] & NF_MISSING) continue;
        LDS uint32_t *h = c->hist + s * 16;
        uint32_t n[16];
#pragma unroll
        for (int i = 0; i < 16; i++) { n[i] = h[i]; h[i] = 0; }
#pragma unroll 1
        for (int k = 0; k < 3; k++) {
            const GBL double *L = c->logt + ((size_t)k * N + s) * 16;
            double e = 0;
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int i = 0; i < 4; i++)
                    if (n[j * 4 + i]) e += (double)n[j * 4 + i] * L[i * 4 + j];
            c->E[s * 3 + k] = e;
        }
        if (c->has_amb && s < c->S) {
            GBL uint32_t *a = c->amb + s * 24;
            for (int k = 0; k < 3; k++) {
                const GBL double *P = c->plin + ((size_t)k * N + s) * 16;
                double e = 0;
                for (int j = 0; j < 4; j++)
                    for (int x = 0; x < 6; x++) {
                        const uint32_t cnt = a[j * 6 + x];
                        if (!cnt) continue;
                        const int code = 6 + x;
                        const int b1 = (code == 6 || code == 9 || code == 11) ? 0 : (code == 7 || code == 10) ? 1 : 2;
                        const int b2 = (code == 6 || code == 10) ? 2 : (code == 9) ? 1 : 3;
                        e += (double)cnt * log(P[j * 4 + b1] + P[j * 4 + b2]);
                    }
                c->E[s * 3 + k] += e;
            }
            for (int i = 0; i < 24; i++) a[i] = 0;
        }
    }
}

// ----------------------------------------------------------------- Z update

// log sum_i exp(T[i*3+j] + y[i]) for parent class j, with the column's own
// maximum taken out first (log_multi2, src/PhyloAcc-ST/bpp.hpp:447-460);
// -inf when every term is -inf.
__device__ __forceinline__ double log_col(const LDS double *T, int j, double y0, double y1, double y2)
{
    const double a = T[j] + y0, b = T[3 + j] + y1, d = T[6 + j] + y2;
    const double mx = fmax(a, fmax(b, d));
    if (mx == -INFINITY) return -INFINITY;
    return mx + log(exp(a - mx) + exp(b - mx) + exp(d - mx));
}

// BPP_C::Update_Z_subtree (src/PhyloAcc-ST/bpp_c.cpp:1182-1293) in log space,
// one tree level per step: upward messages by height, then top-down draws by
// depth.  The draw for node s uses Z-stream slot tz[s].  Log space matters:
// over thousands of sites the classes' emissions differ by thousands of
// nats, and products of linear messages would underflow to zero.
template <typename R, int BS, int SH>
__device__ __noinline__ void update_Z(MP M, CP<R> c, uint32_t c2, uint32_t c3)
{
    const int tid = threadIdx.x;
    const Key key{M->key0, M->key1};
    for (int q = tid; q < c->n_nodes; q += BS) {
        const int s = c->nodes[q];
        double e0, e1, e2;
        if (s == c->root) { e0 = e1 = e2 = 0; }
        else if (c->flags[s] & NF_MISSING) { e0 = log(M->nconstomis); e1 = log(M->constomis); e2 = log(M->nconstomis); }
        else { e0 = c->E[s * 3]; e1 = c->E[s * 3 + 1]; e2 = c->E[s * 3 + 2]; }
        c->beta[s * 3 + 0] = e0;
        c->beta[s * 3 + 1] = e1;
        c->beta[s * 3 + 2] = c->fixZ[s] ? -INFINITY : e2;
        c->changed[s] = 0;
        c->dirty[s] = 0;
    }
    if (tid == 0) c->nch = 0;
    __syncthreads();
    const int n_up = c->n_up, n_down = c->n_down;
    for (int h = 0; h < n_up; h++) {
        // one lane per (node, parent class): the classes' messages are independent
        const int a = h == 0 ? 0 : c->upoff[h - 1], b = c->upoff[h];
        for (int q = 3 * a + tid; q < 3 * b; q += BS) {
            const int p = c->up[q / 3], j = q - 3 * (q / 3);
            double bj = c->beta[p * 3 + j];
            const uint32_t kids = c->desc[4 * (p - c->S)];
            for (int w = 0; w < 2; w++) {
                const int x = w ? (int)(kids >> 16) : (int)(kids & 0xffff);
                const LDS double *T = c->fixZ[x] ? c->TMf : c->TMn;
                bj += log_col(T, j, c->beta[x * 3], c->beta[x * 3 + 1], c->beta[x * 3 + 2]);
            }
            c->beta[p * 3 + j] = bj;
        }
        __syncthreads();
    }
    if (tid == 0) {
        // root prior (1/2, 1/2, 0)
        const int r = c->root;
        const double l0 = c->beta[r * 3] + log(0.5), l1 = c->beta[r * 3 + 1] + log(0.5);
        const double mx = fmax(l0, l1);
        double pr[3] = {exp(l0 - mx), exp(l1 - mx), 0.0};
        const int z = categorical<3>(pr, slot_uniform(key, 0, c2, c3, 0));
        if (c->Z[r] != z) c->changed[r] = 1;
        c->Z[r] = (int8_t)z;
    }
    __syncthreads();
    for (int d = 0; d < n_down; d++) {
        const int a = d == 0 ? 0 : c->downoff[d - 1], b = c->downoff[d];
        for (int q = a + tid; q < b; q += BS) {
            const int s = c->down[q];
            const int zp = c->Z[c->par[s]];
            const LDS double *T = c->fixZ[s] ? c->TMf : c->TMn;
            const double y0 = T[0 * 3 + zp] + c->beta[s * 3], y1 = T[1 * 3 + zp] + c->beta[s * 3 + 1],
                         y2 = T[2 * 3 + zp] + c->beta[s * 3 + 2];
            const double mx = fmax(y0, fmax(y1, y2));
            double pr[3] = {exp(y0 - mx), exp(y1 - mx), exp(y2 - mx)};
            const int z = categorical<3>(pr, slot_uniform(key, 0, c2, c3, (uint32_t)c->tz[s]));
            if (c->Z[s] != z) c->changed[s] = 1;
            c->Z[s] = (int8_t)z;
        }
        __syncthreads();
    }
    // dirty = strict ancestors of changed nodes (getUpdateNode, bpp_c.cpp:620-642)
    const int N = c->N;
    for (int q = tid; q < c->n_nodes; q += BS) {
        const int s = c->nodes[q];
        if (!c->changed[s]) continue;
        c->nch = 1;     // every writer stores the same value
        for (int j = c->par[s]; j != N; j = c->par[j]) c->dirty[j] = 1;
    }
    __syncthreads();
}

// ------------------------------------------------------------- rate update

__device__ __forceinline__ void set_TM(LDS double *F, LDS double *T, double g, double l, double l2)
{
    // log_TM_Int, [i*3+j] = log P(child i | parent j).  fixZ nodes
    // (bpp_c.cpp:305-316, 577-585): column 2 is never set there and keeps its
    // initial log 1 (a fixZ node's parent is never accelerated, so it is unused).
    const double lg = log(g), l1g = log(1 - g);
    F[0] = l1g; F[3] = lg; F[6] = -INFINITY;
    F[1] = -INFINITY; F[4] = 0; F[7] = -INFINITY;
    F[2] = 0; F[5] = 0; F[8] = 0;
    T[0] = l1g; T[3] = lg; T[6] = -INFINITY;
    const double y = 1 - l;   // the reference forms P(1 -> 2) as 1 - (1 - l)
    T[1] = -INFINITY; T[4] = log(y); T[7] = log(1 - y);
    T[2] = -INFINITY; T[5] = log(l2); T[8] = log(1 - l2);
}

// Scalar head of BPP_C::sample_rate (src/PhyloAcc-ST/bpp_c.cpp:930-1041): prior draw
// when no branch carries the class, else the proposal and its range checks.
template <typename R>
__device__ __noinline__ void rate_propose(MP M, CP<R> c, int ms, int m, int resZ, bool neut, double old)
{
    const Key key{M->key0, M->key1};
    const uint32_t c3 = tagword(ms, neut ? TAG_R2 : TAG_R1), c2 = (uint32_t)m;
    const double pa = neut ? M->nprior_a : M->cprior_a, pb = neut ? M->nprior_b : M->cprior_b;
    double cur = old;
    c->skip = 1;
    c->accept = 0;
    if (!c->dirty[c->root]) {
        if (resZ != 0) {
            Stream sp{key, 2, c2, c3, 0};
            int capped = 0;
            cur = gamma_draw(sp, pa, pb, &capped);
            if (capped) c->fail |= PA_F_SAMPLER;
            if (neut) {
                if (M->ropt == 1 && cur < M->nlb) cur = old;
                else if (M->ropt == 2 && cur < c->crate) cur = old;
            } else {
                if (M->ropt == 1 && cur > M->cub) cur = old;
                else if (M->ropt == 2 && cur > c->nrate) cur = old;
            }
        }
    } else {
        const double r = slot_uniform(key, 0, c2, c3, 0);
        double prop;
        int skip = 0;
        if (neut) {
            Stream sg{key, 1, c2, c3, 0};
            int capped = 0;
            prop = gamma_draw(sg, cur / c->prop_n, c->prop_n, &capped);
            if (capped) c->fail |= PA_F_SAMPLER;
            if (M->ropt == 1 && prop < M->nlb) skip = 1;
            else if (M->ropt == 2 && prop < c->crate) skip = 1;
        } else {
            const double ru = M->r1_shared ? r : slot_uniform(key, 1, c2, c3, 0);
            const double u = ru * (c->prop_c - 1 / c->prop_c) + 1 / c->prop_c;
            prop = cur * u;
            if (M->ropt == 1 && prop > M->cub) skip = 1;
            else if (M->ropt == 2 && prop > c->nrate) skip = 1;
        }
        if (!skip && (prop < 1e-12 || prop > 1e8)) {
            skip = 1;
            c->fail |= PA_F_PROPOSAL_RANGE;   // the reference's out-of-range warning (bpp_c.cpp:1024-1041)
        }
        c->skip = skip;
        c->proposal = prop;
        c->r = r;
    }
    if (neut) c->nrate = cur; else c->crate = cur;
}

// Metropolis decision (bpp_c.cpp:1046-1094), FP64.
template <typename R>
__device__ __noinline__ void rate_decide(MP M, CP<R> c, bool neut, double old, double lnew)
{
    const double pa = neut ? M->nprior_a : M->cprior_a, pb = neut ? M->nprior_b : M->cprior_b;
    const double prop = c->proposal;
    double MH;
    if (neut)
        MH = lnew - c->loglik + log_gamma_pdf(prop, pa, pb) - log_gamma_pdf(old, pa, pb) +
             log_gamma_pdf(old, prop / c->prop_n, c->prop_n) - log_gamma_pdf(prop, old / c->prop_n, c->prop_n);
    else
        MH = lnew - c->loglik + log_gamma_pdf(prop, pa, pb) - log_gamma_pdf(old, pa, pb) + log(old) - log(prop);
    if (log(c->r) < MH) {
        c->accept = 1;
        c->loglik = lnew;
        if (neut) c->nrate = prop; else c->crate = prop;
    }
}

// Adaptive proposal scale (bpp_c.cpp:1144-1171).
template <typename R>
__device__ __noinline__ void rate_adapt(MP M, CP<R> c, int m, bool neut)
{
    if ((m + 1) % M->adapt_freq != 0) return;
    double fac = 1 / sqrt((double)((m + 1) / M->adapt_freq));
    if (fac > 0.5) fac = 0.5;
    if (neut) {
        const double ar = (double)c->accept_n / M->adapt_freq;
        const double adj = ar > 0.44 ? exp(fac) : (ar < 0.23 ? exp(-fac) : 1);
        c->prop_n = c->prop_n * adj;
        if (c->prop_n > 1) c->prop_n = 1;
        c->accept_n = 0;
    } else {
        const double ar = (double)c->accept_c / M->adapt_freq;
        const double adj = ar > 0.44 ? exp(fac) : (ar < 0.23 ? exp(-fac) : 1);
        c->prop_c = pow(c->prop_c, adj);
        if (c->prop_c < 1.0001) c->prop_c = 1.0001;
        c->accept_c = 0;
    }
}

template <typename R, int BS, int SH>
__device__ __noinline__ void build_prop_tables(MP M, CP<R> c, int cls, double prop)
{
    for (int q = threadIdx.x; q < c->n_nodes - 1; q += BS) {
        const int s = c->nodes[q];
        if (c->Z[s] != cls || (c->flags[s] & NF_MISSING)) continue;
        const double t = c->dist2[s] * prop;
        double e[4];
#pragma unroll
        for (int k = 0; k < 4; k++) e[k] = exp(M->lam[k] * t);
        LDS R *dst = c->PTp + 16 * s;
#pragma unroll 1
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                double v = 0;
#pragma unroll
                for (int k = 0; k < 4; k++) v += M->V[j * 4 + k] * e[k] * M->Vi[k * 4 + i];
                dst[i * 4 + j] = (R)(v < 0 ? 0.0 : v);
            }
    }
}

template <typename R, int BS, int SH>
__device__ __noinline__ void mark_class(CP<R> c, int cls)
{
    const int tid = threadIdx.x, N = c->N;
    for (int q = tid; q < c->n_nodes; q += BS) c->dirty[c->nodes[q]] = 0;
    __syncthreads();
    for (int q = tid; q < c->n_nodes; q += BS) {
        const int s = c->nodes[q];
        if ((c->flags[s] & NF_MISSING) || c->Z[s] != cls) continue;
        for (int j = c->par[s]; j != N; j = c->par[j]) c->dirty[j] = 1;
    }
}

// BPP_C::getUpdateNode(neut) + BPP_C::sample_rate (src/PhyloAcc-ST/bpp_c.cpp:836-858, 930-1176)
template <typename R, int BS, int SH>
__device__ __forceinline__ void sample_rate(MP M, CP<R> c, int ms, int m, int resZ, bool neut)
{
    const int tid = threadIdx.x;
    const int cls = neut ? 2 : 1;
    const double old = neut ? c->nrate : c->crate;
    mark_class<R, BS, SH>(c, cls);
    __syncthreads();
    if (tid == 0) rate_propose<R>(M, c, ms, m, resZ, neut, old);
    __syncthreads();
    if (!c->skip) {
        build_prop_tables<R, BS, SH>(M, c, cls, c->proposal);
        __syncthreads();
        const double lnew = prune<R, BS, SH>(M, c, PR_PROP, cls);
        if (tid == 0) rate_decide<R>(M, c, neut, old, lnew);
        __syncthreads();
        if (c->accept) {
            for (int q = tid; q < c->n_int; q += BS) {
                const int s = c->inodes[q];
                if (c->dirty[s] && !(c->flags[s] & NF_MISSING)) c->cur[s] ^= 1;
            }
        }
    }
    const double now = neut ? c->nrate : c->crate;
    if (now != old) {
        for (int q = tid; q < c->n_nodes - 1; q += BS) class_table<R>(M, c, cls, c->nodes[q], now, 1);
        __syncthreads();
        refresh_PTc<R, BS, SH>(c, cls);
    }
    __syncthreads();
    if (tid == 0) {
        if (now != old) { if (neut) c->accept_n += 1; else c->accept_c += 1; }
        rate_adapt<R>(M, c, m, neut);
    }
    __syncthreads();
}

// The three Beta draws of BPP_C::sample_transition (src/PhyloAcc-ST/bpp_c.cpp:522-613),
// Beta(a, b) = X / (X + Y) with X ~ Gamma(a) on sub-stream 2k and Y ~ Gamma(b) on
// sub-stream 2k + 1 (beta_draw).  The six gamma variates are independent, so lane j < 6
// draws the one of sub-stream j into red[j]; the l2 pair is not drawn when its prior is off.
template <typename R>
__device__ __noinline__ void transition_gamma(MP M, CP<R> c, int ms, int m, int j)
{
    const Key key{M->key0, M->key1};
    const uint32_t c3 = tagword(ms, TAG_TRANS), c2 = (uint32_t)m;
    double n[9], nf[9];
    for (int i = 0; i < 9; i++) { n[i] = c->cnt9[i]; nf[i] = n[i] - c->cnt9[9 + i]; }
    double shape;
    switch (j) {
    case 0: shape = M->ha_g + n[1 * 3 + 0] + n[2 * 3 + 0]; break;
    case 1: shape = M->hb_g + n[0]; break;
    case 2: shape = M->ha_l + nf[2 * 3 + 1]; break;
    case 3: shape = M->hb_l + nf[1 * 3 + 1]; break;
    case 4: shape = M->ha_l2 + nf[1 * 3 + 2]; break;
    default: shape = M->hb_l2 + nf[2 * 3 + 2]; break;
    }
    double x = 0.0;
    if (j < 4 || M->ha_l2 != 0) {
        Stream st{key, (uint32_t)j, c2, c3, 0};
        int capped = 0;
        x = gamma_draw(st, shape, 1.0, &capped);
        if (capped)   // up to six lanes draw at once
            __hip_atomic_fetch_or(&c->fail, PA_F_SAMPLER, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    c->red[j] = x;
}

template <typename R, int BS, int SH>
__device__ __forceinline__ void sample_transition(MP M, CP<R> c, int ms, int m)
{
    const int tid = threadIdx.x;
    if (tid < 18) c->cnt9[tid] = 0;
    __syncthreads();
    for (int q = tid; q < c->n_nodes - 1; q += BS) {
        const int s = c->nodes[q];
        const int zi = c->Z[s], zp = c->Z[c->par[s]];
        lds_add(&c->cnt9[zi * 3 + zp], 1u);
        if (c->fixZ[s]) lds_add(&c->cnt9[9 + zi * 3 + zp], 1u);
    }
    __syncthreads();
    if (tid < 6) transition_gamma<R>(M, c, ms, m, tid);
    __syncthreads();
    if (tid == 0) {
        const double g = c->red[0] / (c->red[0] + c->red[1]);
        const double l = c->red[2] / (c->red[2] + c->red[3]);
        const double l2 = M->ha_l2 == 0 ? 0.0 : c->red[4] / (c->red[4] + c->red[5]);
        c->g = g; c->l = l; c->l2 = l2;
        set_TM(c->TMf, c->TMn, g, l, l2);
    }
    __syncthreads();
}

// ---------------------------------------------------------- element setup

// A cherry: a non-root internal node whose children are both leaves (tab: the
// element's node table in global memory).
__device__ __forceinline__ bool is_cherry(const int32_t *tab, int N, int S, int root, int x)
{
    return x >= S && x != root && tab[tab_ch0(N) + x] < S && tab[tab_ch1(N) + x] < S;
}

// The static pruning descriptor of internal node s (layout at DB_*): its
// children, each child's kind and missing bit, a cherry child's two leaves
// and their missing bits, and the node's own missing bit.
__device__ __noinline__ void node_desc(LDS uint32_t *d, const int32_t *tab, int N, int S, int root, int s)
{
    const int32_t *ch0 = tab + tab_ch0(N), *ch1 = tab + tab_ch1(N), *fl = tab + tab_flags(N);
    const int x[2] = {ch0[s], ch1[s]};
    uint32_t w[2] = {0, 0};
    uint32_t bits = (fl[s] & NF_MISSING) ? DB_MISS : 0;
    for (int j = 0; j < 2; j++) {
        const int xj = x[j];
        if (xj < 0) continue;
        int kind = KD_STORED;
        if (xj < S) {
            kind = KD_LEAF;
        } else if (is_cherry(tab, N, S, root, xj)) {
            kind = KD_CHERRY;
            const int yA = ch0[xj], yB = ch1[xj];
            w[j] = (uint32_t)yA | ((uint32_t)yB << 16);
            if (fl[yA] & NF_MISSING) bits |= j ? DB_MISS_A1 : DB_MISS_A0;
            if (fl[yB] & NF_MISSING) bits |= j ? DB_MISS_B1 : DB_MISS_B0;
        }
        bits |= (uint32_t)kind << (j ? DB_KIND1_SHIFT : 0);
        if (fl[xj] & NF_MISSING) bits |= j ? DB_MISS1 : DB_MISS0;
    }
    d[0] = (uint32_t)(x[0] < 0 ? 0 : x[0]) | ((uint32_t)(x[1] < 0 ? 0 : x[1]) << 16);
    d[1] = w[0];
    d[2] = w[1];
    d[3] = bits;
}

template <typename R, int BS, int SH>
__device__ __noinline__ void bind_element(MP M, CP<R> c, const ChainArgs *A, int work)
{
    const int tid = threadIdx.x, N = M->N;
    const ElemDev E = A->elems[A->work[work]];
    if (tid == 0) {
        c->S = M->S; c->N = N; c->I = M->I;
        c->GG = E.GG; c->P = E.P; c->Ppad = E.Ppad; c->root = E.root;
        c->n_nodes = E.n_nodes; c->n_int = E.n_int; c->n_up = E.n_up; c->n_down = E.n_down; c->has_amb = E.has_amb;
        c->codes = (const GBL uint8_t *)(A->codes + E.off_codes);
        c->wts = (const GBL int32_t *)(A->wts + E.off_w);
        c->site = (const GBL int32_t *)(A->site + E.off_site);
        char *base = A->scratch + (size_t)blockIdx.x * A->slot_bytes;
        const int64_t bo = A->big_off ? A->big_off[work] : -1;
        char *pbase = bo >= 0 ? A->big + bo : base + A->off_parts;
        c->parts = (GBL V4<R> *)pbase;
        c->exps = (GBL int32_t *)(pbase + chain_exps_off(N, M->S, E.Ppad, sizeof(R)));
        c->plin = (GBL double *)(base + A->off_plin);
        c->logt = (GBL double *)(base + A->off_logt);
        c->amb = (GBL uint32_t *)(base + A->off_amb);
        c->dist2 = (const GBL double *)(A->tab + E.off_tab + tab_dist(N));
        c->accept_n = 0;
        c->accept_c = 0;
        c->bytes = 0;
    }
    const int32_t *tab = A->tab + E.off_tab;
    for (int i = tid; i < N; i += BS) {
        c->par[i] = (int16_t)tab[tab_par(N) + i];
        c->flags[i] = (uint8_t)tab[tab_flags(N) + i];
        c->nodes[i] = (int16_t)tab[tab_nodes(N) + i];
        c->inodes[i] = (int16_t)tab[tab_int(N) + i];
        c->up[i] = (int16_t)tab[tab_up(N) + i];
        c->down[i] = (int16_t)tab[tab_down(N) + i];
        c->tz[i] = (int16_t)tab[tab_tz(N) + i];
        for (int k = 0; k < 16; k++) c->hist[16 * i + k] = 0;
        c->cur[i] = 0;
        c->dirty[i] = 0;
    }
    for (int i = tid; i <= N; i += BS) {
        c->upoff[i] = (int16_t)tab[tab_upoff(N) + i];
        c->downoff[i] = (int16_t)tab[tab_downoff(N) + i];
    }
    for (int s = M->S + tid; s < N; s += BS) node_desc(c->desc + 4 * (s - M->S), tab, N, M->S, E.root, s);
    __syncthreads();
    for (int i = tid; i < M->S * 24; i += BS) c->amb[i] = 0;
    // cherries (both children leaves; never the root) are recomputed where read
    for (int q = tid; q < c->n_int; q += BS) {
        const int s = c->inodes[q];
        if (is_cherry(tab, N, M->S, E.root, s)) c->flags[s] |= NF_CHERRY;
    }
    // null-rate tables (rate 1), fixed for the element
    for (int q = tid; q < c->n_nodes - 1; q += BS) class_table<R>(M, c, 0, c->nodes[q], 1.0, 0);
    __syncthreads();
}

// ------------------------------------------------------------- the chain

template <typename R, int BS, int SH>
__device__ __noinline__ void init_model(MP M, CP<R> c, int resZ)
{
    // BPP_C::initMCMC (src/PhyloAcc-ST/bpp_c.cpp:131-343), iter = 0
    const int tid = threadIdx.x, N = M->N;
    for (int i = tid; i < N; i += BS) {
        c->Z[i] = -1;
        c->maxZ[i] = 0;
    }
    __syncthreads();
    const int root = c->root;
    for (int q = tid; q < c->n_nodes; q += BS) {
        const int s = c->nodes[q];
        const uint8_t f = c->flags[s];
        c->Z[s] = (s == root) ? 0 : 1;
        int fz = (f & NF_UPPER) || s == root;
        if (resZ == 2 && (f & NF_UPPER_CONS)) fz = 1;
        if (resZ == 0 && s != root) fz = 1;
        c->fixZ[s] = (uint8_t)fz;
        c->cur[s] = 0;
    }
    if (tid == 0) {
        c->prop_c = 1.2;
        c->prop_n = (double)1 / c->GG * 10;
        c->max_loglik = -INFINITY;
        c->max_m = 0;
        c->nrate = M->init_nrate;
        c->crate = M->init_crate;
        c->g = M->init_grate; c->l = M->init_lrate; c->l2 = M->ha_l2 == 0 ? 0.0 : M->init_lrate2;
        set_TM(c->TMf, c->TMn, c->g, c->l, c->l2);
        c->fail = 0;
    }
    __syncthreads();
    const double cr = c->crate, nr = c->nrate;
    for (int q = tid; q < c->n_nodes - 1; q += BS) {
        const int s = c->nodes[q];
        class_table<R>(M, c, 1, s, cr, 0);
        class_table<R>(M, c, 2, s, nr, 0);
    }
    __syncthreads();
    refresh_PTc<R, BS, SH>(c, -1);
    __syncthreads();
}

// BPP_C::MonitorChain (src/PhyloAcc-ST/bpp_c.cpp:741-834), thread 0
template <typename R>
__device__ __noinline__ void monitor(MP M, CP<R> c, int m, int resZ, double *trd, int T)
{
    double full = c->loglik + log_gamma_pdf(c->crate, M->cprior_a, M->cprior_b);
    if (resZ != 0) full += log_gamma_pdf(c->nrate, M->nprior_a, M->nprior_b);
    if (trd) {
        trd[0 * T + m] = c->loglik; trd[1 * T + m] = full; trd[2 * T + m] = c->nrate;
        trd[3 * T + m] = c->crate; trd[4 * T + m] = c->g; trd[5 * T + m] = c->l; trd[6 * T + m] = c->l2;
    }
    if (m >= M->burnin && full > c->max_loglik) {
        c->max_loglik = full;
        c->max_m = m;
        for (int i = 0; i < M->N; i++) c->maxZ[i] = c->Z[i];
    }
    if (isnan(c->loglik)) c->fail |= PA_F_NAN;
    else if (isinf(c->loglik)) c->fail |= c->loglik > 0 ? PA_F_POS_INF : PA_F_NEG_INF;
}

template <typename R, int BS, int SH>
__device__ __noinline__ void write_results(MP M, CP<R> c, const ChainArgs *A, int ms, int work)
{
    const int tid = threadIdx.x, N = M->N;
    if (tid == 0) {
        double *rd = A->res_d + ((size_t)work * 3 + ms) * RES_D;
        int32_t *ri = A->res_i + ((size_t)work * 3 + ms) * RES_I;
        rd[0] = c->max_loglik; rd[1] = c->nrate; rd[2] = c->crate; rd[3] = c->g; rd[4] = c->l; rd[5] = c->l2;
        const bool failed = (c->fail & (PA_F_NAN | PA_F_POS_INF | PA_F_SAMPLER)) || !isfinite(c->max_loglik);
        ri[0] = c->max_m; ri[1] = failed ? PA_STATUS_NUMERIC_FAIL : 0; ri[2] = c->accept_n; ri[3] = c->accept_c;
        ri[4] = c->fail;
    }
    int8_t *mz = A->maxZ + ((size_t)work * 3 + ms) * N;
    for (int i = tid; i < N; i += BS) mz[i] = c->maxZ[i];
}

template <typename R, int BS, int SH>
__device__ __forceinline__ void run_model(MP M, CP<R> c, const ChainArgs *A, int ms, int work)
{
    const int tid = threadIdx.x;
    const int resZ = ms == 0 ? 0 : (ms == 1 ? 2 : 1);
    const int T = M->T, N = M->N;
    init_model<R, BS, SH>(M, c, resZ);
    {
        const double l0 = prune<R, BS, SH>(M, c, PR_FULL, 0);
        if (tid == 0) c->loglik = l0;
        __syncthreads();
    }
    double *trd = A->tr_d ? A->tr_d + ((size_t)work * 3 + ms) * 7 * T : nullptr;
    int8_t *trz = A->tr_Z ? A->tr_Z + ((size_t)work * 3 + ms) * (size_t)T * N : nullptr;
    for (int m = 0; m < T; m++) {
        if (trz)
            for (int i = tid; i < N; i += BS) trz[(size_t)m * N + i] = c->Z[i];
        if (tid == 0) monitor<R>(M, c, m, resZ, trd, T);
        if (m == T - 1) break;
        sample_sites<R, BS, SH>(M, c, (uint32_t)m, tagword(ms, TAG_ANC));
        __syncthreads();
        emissions<R, BS, SH>(c);
        __syncthreads();
        update_Z<R, BS, SH>(M, c, (uint32_t)m, tagword(ms, TAG_Z));
        refresh_PTc<R, BS, SH>(c, REFRESH_CHANGED);
        __syncthreads();
        if (c->nch) {
            const double l = prune<R, BS, SH>(M, c, PR_DIRTY, 0);
            if (tid == 0) c->loglik = l;
            __syncthreads();
        }
        sample_rate<R, BS, SH>(M, c, ms, m, resZ, true);
        sample_rate<R, BS, SH>(M, c, ms, m, resZ, false);
        sample_transition<R, BS, SH>(M, c, ms, m);
    }
    __syncthreads();
    write_results<R, BS, SH>(M, c, A, ms, work);
    __syncthreads();
}

__device__ __forceinline__ MP model_to_lds(const ModelDev &Mk, LDS ModelDev *Ms)
{
    const uint32_t *src = (const uint32_t *)&Mk;
    LDS uint32_t *dst = (LDS uint32_t *)Ms;
    for (int i = threadIdx.x; i < (int)(sizeof(ModelDev) / 4); i += blockDim.x) dst[i] = src[i];
    return (MP)Ms;
}

template <typename R, int BS, int SH>
__global__ void __launch_bounds__(BS, 4) chain_kernel(ModelDev Mk, ChainArgs A)
{
    extern __shared__ __align__(16) char smem_raw[];
    __shared__ ModelDev Ms;
    __shared__ Ctx<R> Cs;
    MP M = model_to_lds(Mk, (LDS ModelDev *)&Ms);
    CP<R> c = (CP<R>)&Cs;
    __syncthreads();
    if (threadIdx.x == 0) smem_arrays<R, true>(M->N, M->S, BS, (LDS char *)smem_raw, c);
    __syncthreads();
    for (;;) {
        if (threadIdx.x == 0) c->work = atomicAdd(A.queue, 1);
        __syncthreads();
        const int w = c->work;
        __syncthreads();
        if (w >= A.n_work) break;
        bind_element<R, BS, SH>(M, c, &A, w);
        for (int ms = 0; ms < 3; ms++) run_model<R, BS, SH>(M, c, &A, ms, w);
        if (threadIdx.x == 0 && A.counters) atomicAdd(A.counters, c->bytes);
        if (threadIdx.x == 0 && A.done) {
            // thread 0 wrote the result rows: make them visible to the host
            // before the flag that announces them
            __threadfence_system();
            __hip_atomic_store(A.done + w, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

__global__ void reset_queue(int *q) { *q = 0; }

template <typename R, int BS, int SH>
static hipError_t launch_t(const ModelDev &M, const ChainArgs &A, int grid, hipStream_t st)
{
    const size_t sm = smem_arrays<R, false>(M.N, M.S, BS, nullptr, nullptr);
    const hipError_t err = hipFuncSetAttribute((const void *)chain_kernel<R, BS, SH>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL((chain_kernel<R, BS, SH>), dim3(grid), dim3(BS), sm, st, M, A);
    return hipGetLastError();
}

hipError_t launch_reset(int *q, hipStream_t st)
{
    hipLaunchKernelGGL(reset_queue, dim3(1), dim3(1), 0, st, q);
    return hipGetLastError();
}

// Zeroes n 32-bit words: a kernel, so that it can be captured in a HIP graph
// where a memset node would not be used.
__global__ void clear_words(uint32_t *p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

hipError_t launch_clear(void *p, int words, hipStream_t st)
{
    hipLaunchKernelGGL(clear_words, dim3(1), dim3(64), 0, st, (uint32_t *)p, words);
    return hipGetLastError();
}

template <typename R, int BS, int SH>
static int blocks_t(int N, int S)
{
    const size_t sm = smem_arrays<R, false>(N, S, BS, nullptr, nullptr);
    if (hipFuncSetAttribute((const void *)chain_kernel<R, BS, SH>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm) != hipSuccess)
        return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, chain_kernel<R, BS, SH>, BS, sm) != hipSuccess) return 0;
    return nb;
}

// Runtime (precision, block size, shape) -> template instance.
template <typename R, int SH, typename F>
static auto by_block(int BS, F &&f) -> decltype(f(std::integral_constant<int, 64>()))
{
    if (BS == 64) return f(std::integral_constant<int, 64>());
    if (BS == 128) return f(std::integral_constant<int, 128>());
    return f(std::integral_constant<int, 256>());
}

template <typename R, typename F>
static auto by_shape(int shape, F &&f) -> decltype(f(std::integral_constant<int, 0>()))
{
    if (shape == 1) return f(std::integral_constant<int, 1>());
    if (shape == 2) return f(std::integral_constant<int, 2>());
    return f(std::integral_constant<int, 0>());
}

template <typename R>
static int blocks_r(int BS, int shape, int N, int S)
{
    return by_shape<R>(shape, [&](auto sh) {
        return by_block<R, decltype(sh)::value>(BS, [&](auto bs) { return blocks_t<R, decltype(bs)::value, decltype(sh)::value>(N, S); });
    });
}

template <typename R>
static hipError_t launch_r(const ModelDev &M, const ChainArgs &A, int BS, int shape, int grid, hipStream_t st)
{
    return by_shape<R>(shape, [&](auto sh) {
        return by_block<R, decltype(sh)::value>(BS, [&](auto bs) {
            return launch_t<R, decltype(bs)::value, decltype(sh)::value>(M, A, grid, st);
        });
    });
}

int chain_blocks_per_cu(int fp64, int BS, int N, int S, int shape)
{
    if ((BS != 64 && BS != 128 && BS != 256) || shape < 0 || shape >= N_SHAPES) return 0;
    return fp64 ? blocks_r<double>(BS, shape, N, S) : blocks_r<float>(BS, shape, N, S);
}

hipError_t launch_chain(const ModelDev &M, const ChainArgs &A, int fp64, int BS, int grid, hipStream_t st, int shape)
{
    if ((BS != 64 && BS != 128 && BS != 256) || shape < 0 || shape >= N_SHAPES) return hipErrorInvalidValue;
    return fp64 ? launch_r<double>(M, A, BS, shape, grid, st) : launch_r<float>(M, A, BS, shape, grid, st);
}

}  // namespace chain
}  // namespace pa
