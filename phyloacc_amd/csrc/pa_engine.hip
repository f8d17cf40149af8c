// pa_engine.hip -- CDNA4 (gfx950) kernels for PhyloAcc-ST's per-element MCMC.
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
//   * per-branch 4x4 transition matrices of the current rate class live in
//     LDS (read as broadcasts); FP64 is used for P(t) builds, log tables,
//     emissions, the Z forward-backward, likelihood sums and MH ratios.
// The chain context and the model constants live in LDS behind
// address-space-3 pointers, and every sweep phase is a separate non-inlined
// function, so each phase gets its own register allocation instead of one
// allocation spanning the whole persistent loop.
// No MFMA: a K=4 matvec per pattern is not a dense contraction.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <stdint.h>
#include "pa_common.h"
#include "pa_rng.h"
#include "pa_kernels.h"

#define LDS __attribute__((address_space(3)))
#define GBL __attribute__((address_space(1)))

namespace pa {

template <typename R>
struct alignas(4 * sizeof(R)) V4 { R x, y, z, w; };

// explicit member-wise loads/stores through address-space pointers
template <typename R, typename P>
__device__ __forceinline__ V4<R> ld4(const P *p)
{
    V4<R> v;
    v.x = p->x; v.y = p->y; v.z = p->z; v.w = p->w;
    return v;
}
template <typename R, typename P>
__device__ __forceinline__ void st4(P *p, R a, R b, R c, R d)
{
    p->x = a; p->y = b; p->z = c; p->w = d;
}

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
    // LDS arrays
    LDS int16_t *ch0, *ch1, *par, *nodes, *inodes, *up, *upoff, *down, *downoff, *tz, *lvl, *kbf;
    LDS uint8_t *flags, *fixZ, *dirty, *cur, *changed;
    LDS int8_t *Z, *maxZ;
    LDS double *dist2, *E, *beta;
    LDS R *PTc, *PTp;      // [N][16], PT[i*4+j] = P_{j->i}
    LDS uint32_t *hist, *cnt9;
    LDS uint8_t *tgs;
    // chain scalars
    double loglik, nrate, crate, g, l, l2, prop_n, prop_c, proposal, r, max_loglik;
    double TMf[9], TMn[9];  // linear log_TM_Int for fixZ / free nodes, [i*3+j] = P(child i | parent j)
    double red[16];
    unsigned long long bytes;   // algorithmic HBM bytes of this element's passes (see prune / sample_sites)
    int accept_n, accept_c, max_m, skip, accept, status, work, nch;
};

template <typename R>
using CP = LDS Ctx<R> *;
using MP = LDS const ModelDev *;
using AP = LDS const ChainArgs *;

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
        o = al(o, 32);  /* rows of PT are read as 4-vectors */                      \
        if constexpr (BIND) c->field = (LDS T *)(base + o);                        \
        o += sizeof(T) * (size_t)(count);                                          \
    } while (0)
    PA_TAKE(dist2, double, N);
    PA_TAKE(PTc, R, 16 * N);
    // Phases that never overlap share storage:
    //   A: per-thread ancestral states (sampling) | emissions + Z messages;
    //   B: per-branch transition counters (sampling, emissions) | proposal
    //      matrices (rate updates).  The counters are cleared when a sweep's
    //      sampling starts.
    {
        const size_t szA = std::max((size_t)(I > 0 ? I : 1) * BS, (size_t)6 * N * sizeof(double));
        o = al(o, 32);
        if constexpr (BIND) {
            c->tgs = (LDS uint8_t *)(base + o);
            c->E = (LDS double *)(base + o);
            c->beta = (LDS double *)(base + o) + 3 * N;
        }
        o += szA;
        const size_t szB = std::max((size_t)16 * N * sizeof(uint32_t), (size_t)16 * N * sizeof(R));
        o = al(o, 32);
        if constexpr (BIND) {
            c->hist = (LDS uint32_t *)(base + o);
            c->PTp = (LDS R *)(base + o);
        }
        o += szB;
    }
    PA_TAKE(cnt9, uint32_t, 20);
    PA_TAKE(ch0, int16_t, N);
    PA_TAKE(ch1, int16_t, N);
    PA_TAKE(par, int16_t, N);
    PA_TAKE(nodes, int16_t, N);
    PA_TAKE(inodes, int16_t, N);
    PA_TAKE(up, int16_t, N);
    PA_TAKE(upoff, int16_t, N + 1);
    PA_TAKE(down, int16_t, N);
    PA_TAKE(downoff, int16_t, N + 1);
    PA_TAKE(tz, int16_t, N);
    PA_TAKE(lvl, int16_t, N);
    PA_TAKE(kbf, int16_t, N + 1);
    PA_TAKE(flags, uint8_t, N);
    PA_TAKE(fixZ, uint8_t, N);
    PA_TAKE(dirty, uint8_t, N);
    PA_TAKE(cur, uint8_t, N);
    PA_TAKE(changed, uint8_t, N);
    PA_TAKE(Z, int8_t, N);
    PA_TAKE(maxZ, int8_t, N);
#undef PA_TAKE
    return al(o, 16);
}

size_t chain_smem_bytes(int N, int S, int BS, int fp64)
{
    return (fp64 ? smem_arrays<double, false>(N, S, BS, nullptr, nullptr)
                 : smem_arrays<float, false>(N, S, BS, nullptr, nullptr)) +
           (fp64 ? sizeof(Ctx<double>) : sizeof(Ctx<float>)) + sizeof(ModelDev) + sizeof(ChainArgs);
}

// ------------------------------------------------------------------ helpers

template <typename R>
__device__ __forceinline__ double block_sum(double v, CP<R> c)
{
    // deterministic: fixed xor tree inside each wave, waves summed in order
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) c->red[w] = v;
    __syncthreads();
    double t = 0;
    for (int i = 0; i < nw; i++) t += c->red[i];
    __syncthreads();
    return t;
}

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
            out[j * 4 + i] = s > 0 ? s : 0.0;
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

// PT of the current class of every branch (all branches, or those with Z == only_cls)
template <typename R, int BS>
__device__ __noinline__ void refresh_PTc(CP<R> c, int only_cls)
{
    for (int q = threadIdx.x; q < c->n_nodes - 1; q += BS) {
        const int s = c->nodes[q];
        const int z = c->Z[s];
        if (only_cls >= 0 && z != only_cls) continue;
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

template <typename R>
struct M16 { R v[16]; };

template <typename R>
__device__ __forceinline__ void leaf_mul(const LDS R *PT, int code, R &a0, R &a1, R &a2, R &a3)
{
    if (code < 4) {
        const V4<R> row = ld4<R>((const LDS V4<R> *)(PT + code * 4));
        a0 *= row.x; a1 *= row.y; a2 *= row.z; a3 *= row.w;
    } else {
        // ambiguity codes r y k m s w -> base pairs (a,g) (c,t) (g,t) (a,c) (c,g) (a,t)
        const int b1 = (code == 6 || code == 9 || code == 11) ? 0 : (code == 7 || code == 10) ? 1 : 2;
        const int b2 = (code == 6 || code == 10) ? 2 : (code == 9) ? 1 : 3;
        const V4<R> r1 = ld4<R>((const LDS V4<R> *)(PT + b1 * 4));
        const V4<R> r2 = ld4<R>((const LDS V4<R> *)(PT + b2 * 4));
        a0 *= r1.x + r2.x; a1 *= r1.y + r2.y; a2 *= r1.z + r2.z; a3 *= r1.w + r2.w;
    }
}

// One child's inputs at one pattern: a leaf's code, or an internal child's
// scaled partial and cumulative exponent.
template <typename R>
struct ChildIn {
    V4<R> v;
    int x;
};

template <typename R>
__device__ __forceinline__ ChildIn<R> child_load(bool leaf, bool miss, const GBL uint8_t *code_row,
                                                 const GBL V4<R> *part_row, const GBL int32_t *exp_row, int p)
{
    // a missing child contributes nothing, but its gap state still decides
    // whether the parent is all-gap at this pattern, so its code / exponent
    // is read either way
    ChildIn<R> in;
    in.v.x = in.v.y = in.v.z = in.v.w = (R)0;
    if (leaf) {
        in.x = code_row[p];
    } else {
        in.x = exp_row[p];
        if (!miss) in.v = ld4<R>(part_row + p);
    }
    return in;
}

// Multiplies one child's message P * L_child into (a0..a3).  A leaf reads the
// row of its base from the LDS matrix; an internal child uses the register
// copy m.  gap reports the child's GAPT state (Tg >= num_base).
template <typename R>
__device__ __forceinline__ void child_apply(bool leaf, bool miss, const LDS R *src, const M16<R> &m,
                                            const ChildIn<R> &in, R &a0, R &a1, R &a2, R &a3, int &e, bool &gap)
{
    if (leaf) {
        gap = (in.x == PA_GAP_CODE || in.x == PA_OTHER_CODE);
        if (!miss && !gap) leaf_mul<R>(src, in.x, a0, a1, a2, a3);
    } else {
        gap = (in.x == GAPT);
        if (!miss && !gap) {
            const V4<R> v = in.v;
            const R o0 = m.v[0] * v.x + m.v[4] * v.y + m.v[8] * v.z + m.v[12] * v.w;
            const R o1 = m.v[1] * v.x + m.v[5] * v.y + m.v[9] * v.z + m.v[13] * v.w;
            const R o2 = m.v[2] * v.x + m.v[6] * v.y + m.v[10] * v.z + m.v[14] * v.w;
            const R o3 = m.v[3] * v.x + m.v[7] * v.y + m.v[11] * v.z + m.v[15] * v.w;
            a0 *= o0; a1 *= o1; a2 *= o2; a3 *= o3;
            e += in.x;
        }
    }
}

// Values every lane of the workgroup holds alike (node ids, offsets read
// from LDS) are moved to scalar registers, which frees vector registers and
// makes the address arithmetic scalar.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
template <typename T>
__device__ __forceinline__ T *uni_ptr(T *p)
{
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return (T *)(((uint64_t)hi << 32) | lo);
}

// Patterns handled per thread per round: all their loads are issued before
// the first is used, so a round costs one memory latency.
template <typename R>
struct Unroll { static constexpr int U = sizeof(R) == 4 ? 4 : 2; };

// Nodes a pass visits, in level order (height 1, 2, ...): all internal nodes
// (FULL) or the non-missing dirty ones (DIRTY / PROP).  Wave 0 compacts the
// height-ordered list with ballots; lvl[] receives the nodes and kbf[q] the
// number kept before position q of the height-ordered list.
template <typename R, int BS>
__device__ __forceinline__ void visit_list(CP<R> c, int mode)
{
    const int tid = threadIdx.x, n_int = c->n_int;
    if (tid < 64) {
        int kept = 0;
        for (int base = 0; base < n_int; base += 64) {
            const int q = base + tid;
            bool keep = false;
            int s = 0;
            if (q < n_int) {
                s = c->up[q];
                keep = mode == PR_FULL || (c->dirty[s] && !(c->flags[s] & NF_MISSING));
            }
            const unsigned long long bal = __ballot(keep);
            const int before = kept + __popcll(bal & ((1ull << tid) - 1ull));
            if (q < n_int) {
                c->kbf[q] = (int16_t)before;
                if (keep) c->lvl[before] = (int16_t)s;
            }
            kept += __popcll(bal);
        }
        if (tid == 0) c->kbf[n_int] = (int16_t)kept;
    }
    __syncthreads();
}

// One pruning pass (BPP_C::Update_Tg(tosample=false), src/PhyloAcc-ST/bpp_c.cpp:644-671,
// and log_f_Xz, bpp_c.cpp:860-928).  FULL: every internal node of the tree
// (missing nodes only record their GAPT flag) into the current buffer.
// DIRTY: nodes with dirty[] set, in place.  PROP: dirty nodes into the other
// buffer, with PTp for branches whose Z is cls.  Returns the element's
// log-likelihood (FP64), identical on every thread.
//
// A thread owns the patterns p = tid (mod BS) of every node, so it only ever
// reads partials it wrote itself and the pass needs no barrier.  Nodes of one
// tree level are independent: a round takes U (node, pattern block) items of
// the same level and issues all their loads before the first is used, so a
// round costs one memory latency whether the level is wide (many nodes) or
// the element is long (many pattern blocks).
template <typename R, int BS>
__device__ __noinline__ double prune(MP M, CP<R> c, int mode, int cls)
{
    const int tid = threadIdx.x;
    const int P = uni(c->P), Ppad = uni(c->Ppad), S = uni(c->S);
    const size_t bufI = (size_t)uni(c->I) * Ppad;
    GBL V4<R> *parts = uni_ptr(c->parts);
    GBL int32_t *exps = uni_ptr(c->exps);
    const GBL uint8_t *codes = uni_ptr(c->codes);
    visit_list<R, BS>(c, mode);
    const int Pb = uni((P + BS - 1) / BS);
    constexpr int U = Unroll<R>::U;
    unsigned long long bytes = 0;
    for (int h = 0; h < c->n_up; h++) {
        const int a = uni(c->kbf[h == 0 ? 0 : c->upoff[h - 1]]), b = uni(c->kbf[c->upoff[h]]);
        const int items = (b - a) * Pb;
        for (int i0 = 0; i0 < items; i0 += U) {
            ChildIn<R> in0[U], in1[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u;
                if (i >= items) break;
                const int q = i / Pb, p = (i - q * Pb) * BS + tid;
                const int s = uni(c->lvl[a + q]);
                const int x0 = uni(c->ch0[s]), x1 = uni(c->ch1[s]);
                const bool leaf0 = x0 < S, leaf1 = x1 < S;
                const bool miss0 = uni(c->flags[x0]) & NF_MISSING, miss1 = uni(c->flags[x1]) & NF_MISSING;
                const int rb0 = leaf0 ? 0 : ((mode == PR_PROP && c->dirty[x0]) ? 1 - c->cur[x0] : c->cur[x0]);
                const int rb1 = leaf1 ? 0 : ((mode == PR_PROP && c->dirty[x1]) ? 1 - c->cur[x1] : c->cur[x1]);
                const size_t o0 = leaf0 ? 0 : rb0 * bufI + (size_t)(x0 - S) * Ppad;
                const size_t o1 = leaf1 ? 0 : rb1 * bufI + (size_t)(x1 - S) * Ppad;
                if (p < P) {
                    in0[u] = child_load<R>(leaf0, miss0, codes + (size_t)x0 * Ppad, parts + o0, exps + o0, p);
                    in1[u] = child_load<R>(leaf1, miss1, codes + (size_t)x1 * Ppad, parts + o1, exps + o1, p);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                // one item at a time: keeps the scheduler from hoisting every
                // item's matrices into registers at once
                __builtin_amdgcn_sched_barrier(0);
                const int i = i0 + u;
                if (i >= items) break;
                const int q = i / Pb, p = (i - q * Pb) * BS + tid;
                const int s = uni(c->lvl[a + q]);
                const int x0 = uni(c->ch0[s]), x1 = uni(c->ch1[s]);
                const bool leaf0 = x0 < S, leaf1 = x1 < S;
                const bool miss_s = uni(c->flags[s]) & NF_MISSING;
                const bool miss0 = uni(c->flags[x0]) & NF_MISSING, miss1 = uni(c->flags[x1]) & NF_MISSING;
                if (tid == 0 && i - q * Pb == 0)   // write partial + exponent; read each child's partial + exponent, or its leaf code
                    bytes += (unsigned long long)P * ((4 * sizeof(R) + 4) + (leaf0 ? 1 : 4 * sizeof(R) + 4) + (leaf1 ? 1 : 4 * sizeof(R) + 4));
                if (p >= P) continue;
                const LDS R *src0 = (mode == PR_PROP && c->Z[x0] == cls) ? c->PTp + 16 * x0 : c->PTc + 16 * x0;
                const LDS R *src1 = (mode == PR_PROP && c->Z[x1] == cls) ? c->PTp + 16 * x1 : c->PTc + 16 * x1;
                M16<R> m0, m1;
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    m0.v[k] = leaf0 ? (R)0 : src0[k];
                    m1.v[k] = leaf1 ? (R)0 : src1[k];
                }
                const int wb = (mode == PR_PROP) ? 1 - c->cur[s] : c->cur[s];
                GBL V4<R> *dstp = parts + wb * bufI + (size_t)(s - S) * Ppad;
                GBL int32_t *dste = exps + wb * bufI + (size_t)(s - S) * Ppad;
                R a0 = 1, a1 = 1, a2 = 1, a3 = 1;
                int e = 0;
                bool g0, g1;
                child_apply<R>(leaf0, miss0, src0, m0, in0[u], a0, a1, a2, a3, e, g0);
                child_apply<R>(leaf1, miss1, src1, m1, in1[u], a0, a1, a2, a3, e, g1);
                if (g0 && g1) {
                    dste[p] = GAPT;
                } else if (miss_s) {
                    dste[p] = 0;
                } else {
                    const R mx = fmax(fmax(a0, a1), fmax(a2, a3));
                    int kx = 0;
                    frexp(mx, &kx);
                    a0 = ldexp(a0, -kx); a1 = ldexp(a1, -kx); a2 = ldexp(a2, -kx); a3 = ldexp(a3, -kx);
                    st4<R>(dstp + p, a0, a1, a2, a3);
                    dste[p] = e + kx;
                }
            }
        }
    }
    if (tid == 0) c->bytes += bytes;
    // root: log sum_j pi_j L_j + exponent * ln 2; all-gap columns contribute 0
    const int r = c->root;
    double acc = 0;
    if (r >= S) {
        const int rb = (mode == PR_PROP && c->dirty[r]) ? 1 - c->cur[r] : c->cur[r];
        const GBL V4<R> *rp = parts + rb * bufI + (size_t)(r - S) * Ppad;
        const GBL int32_t *re = exps + rb * bufI + (size_t)(r - S) * Ppad;
        const double p0 = M->pi[0], p1 = M->pi[1], p2 = M->pi[2], p3 = M->pi[3];
        if (tid == 0) c->bytes += (unsigned long long)P * (4 * sizeof(R) + 4 + 4);   // root partial, exponent, weight
        for (int pb = tid; pb < P; pb += U * BS) {
            int ex[U], wt[U];
            V4<R> av[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int p = pb + u * BS;
                ex[u] = GAPT;
                if (p < P) {
                    ex[u] = re[p];
                    wt[u] = c->wts[p];
                    av[u] = ld4<R>(rp + p);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (ex[u] == GAPT) continue;
                const double L = p0 * (double)av[u].x + p1 * (double)av[u].y + p2 * (double)av[u].z + p3 * (double)av[u].w;
                acc += (double)wt[u] * (log(L) + (double)ex[u] * 0.69314718055994530942);
            }
        }
    }
    return block_sum<R>(acc, c);
}

// ---------------------------------------------------------- element setup

template <typename R, int BS>
__device__ __noinline__ void bind_element(MP M, CP<R> c, AP A, int eid, int rec)
{
    const int tid = threadIdx.x, N = M->N;
    const ElemDev E = A->elems[eid];
    if (tid == 0) {
        c->S = M->S; c->N = N; c->I = M->I;
        c->GG = E.GG; c->P = E.P; c->Ppad = E.Ppad; c->root = E.root;
        c->n_nodes = E.n_nodes; c->n_int = E.n_int; c->n_up = E.n_up; c->n_down = E.n_down; c->has_amb = E.has_amb;
        c->codes = (const GBL uint8_t *)(A->codes + E.off_codes);
        c->wts = (const GBL int32_t *)(A->wts + E.off_w);
        c->site = (const GBL int32_t *)(A->site + E.off_site);
        char *base = A->scratch + (A->slot_off ? A->slot_off[rec] : (int64_t)blockIdx.x * A->slot_bytes);
        size_t off[5];
        slot_layout(N, M->S, E.Ppad, sizeof(R) == 8, off);
        c->parts = (GBL V4<R> *)(base + off[0]);
        c->exps = (GBL int32_t *)(base + off[1]);
        c->plin = (GBL double *)(base + off[2]);
        c->logt = (GBL double *)(base + off[3]);
        c->amb = (GBL uint32_t *)(base + off[4]);
        c->accept_n = 0;
        c->accept_c = 0;
        c->bytes = 0;
    }
    const int32_t *tab = A->tab + E.off_tab;
    for (int i = tid; i < N; i += BS) {
        c->ch0[i] = (int16_t)tab[tab_ch0(N) + i];
        c->ch1[i] = (int16_t)tab[tab_ch1(N) + i];
        c->par[i] = (int16_t)tab[tab_par(N) + i];
        c->flags[i] = (uint8_t)tab[tab_flags(N) + i];
        c->nodes[i] = (int16_t)tab[tab_nodes(N) + i];
        c->inodes[i] = (int16_t)tab[tab_int(N) + i];
        c->up[i] = (int16_t)tab[tab_up(N) + i];
        c->down[i] = (int16_t)tab[tab_down(N) + i];
        c->tz[i] = (int16_t)tab[tab_tz(N) + i];
        c->dist2[i] = ((const double *)(tab + tab_dist(N)))[i];
        for (int k = 0; k < 16; k++) c->hist[16 * i + k] = 0;
        c->cur[i] = 0;
        c->dirty[i] = 0;
    }
    for (int i = tid; i <= N; i += BS) {
        c->upoff[i] = (int16_t)tab[tab_upoff(N) + i];
        c->downoff[i] = (int16_t)tab[tab_downoff(N) + i];
    }
    __syncthreads();
    for (int i = tid; i < M->S * 24; i += BS) c->amb[i] = 0;
    // null-rate tables (rate 1), fixed for the element
    for (int q = tid; q < c->n_nodes - 1; q += BS) class_table<R>(M, c, 0, c->nodes[q], 1.0, 0);
    __syncthreads();
}

// Kernel arguments are copied to LDS once per workgroup: the device
// functions take LDS pointers to them, which keeps the argument blocks out of
// the private (scratch) segment.
template <typename T>
__device__ __forceinline__ LDS const T *to_lds(const T &k, LDS T *dst)
{
    static_assert(sizeof(T) % 4 == 0, "word copy");
    const uint32_t *src = (const uint32_t *)&k;
    LDS uint32_t *d = (LDS uint32_t *)dst;
    for (int i = threadIdx.x; i < (int)(sizeof(T) / 4); i += blockDim.x) d[i] = src[i];
    return (LDS const T *)dst;
}

// Log-likelihood at fixed Z and rates (BPP_C::log_f_Xz, bpp_c2.cpp:392-425)
template <typename R, int BS>
__global__ void __launch_bounds__(BS) loglik_kernel(ModelDev Mk, ChainArgs Ak)
{
    extern __shared__ __align__(16) char smem_raw[];
    __shared__ ModelDev Ms;
    __shared__ ChainArgs As;
    __shared__ Ctx<R> Cs;
    MP M = to_lds(Mk, (LDS ModelDev *)&Ms);
    AP A = to_lds(Ak, (LDS ChainArgs *)&As);
    CP<R> c = (CP<R>)&Cs;
    __syncthreads();
    if (threadIdx.x == 0) smem_arrays<R, true>(M->N, M->S, BS, (LDS char *)smem_raw, c);
    __syncthreads();
    for (;;) {
        if (threadIdx.x == 0) c->work = atomicAdd(A->queue, 1);
        __syncthreads();
        const int w = c->work;
        __syncthreads();
        if (w >= A->n_work) break;
        bind_element<R, BS>(M, c, A, A->work[w], w);
        const double r1 = A->rates_in[2 * w], r2 = A->rates_in[2 * w + 1];
        for (int i = threadIdx.x; i < M->N; i += BS) c->Z[i] = (int8_t)A->Zin[(size_t)w * M->N + i];
        if (A->len_in) {
            // own-tree element: per-call lengths, every branch in the rate-1 class
            for (int i = threadIdx.x; i < M->N; i += BS) c->dist2[i] = A->len_in[(size_t)w * M->N + i];
            __syncthreads();
            for (int q = threadIdx.x; q < c->n_nodes - 1; q += BS) class_table<R>(M, c, 0, c->nodes[q], 1.0, 0);
        }
        for (int q = threadIdx.x; q < c->n_nodes - 1; q += BS) {
            const int s = c->nodes[q];
            class_table<R>(M, c, 1, s, r1, 0);
            class_table<R>(M, c, 2, s, r2, 0);
        }
        __syncthreads();
        refresh_PTc<R, BS>(c, -1);
        __syncthreads();
        const double l = prune<R, BS>(M, c, PR_FULL, 0);
        if (threadIdx.x == 0) A->ll_out[w] = l;
        __syncthreads();
    }
}

__global__ void reset_queue(int *q) { *q = 0; }

template <typename R, int BS>
static hipError_t launch_t(const ModelDev &M, const ChainArgs &A, int grid, hipStream_t st)
{
    // dynamic LDS above 64 KB needs the per-kernel limit raised (large trees only)
    const size_t sm = smem_arrays<R, false>(M.N, M.S, BS, nullptr, nullptr);
    if (sm > 65536) {
        const hipError_t err = hipFuncSetAttribute((const void *)loglik_kernel<R, BS>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
        if (err != hipSuccess) return err;
    }
    hipLaunchKernelGGL((loglik_kernel<R, BS>), dim3(grid), dim3(BS), sm, st, M, A);
    return hipGetLastError();
}

hipError_t launch_reset(int *q, hipStream_t st)
{
    hipLaunchKernelGGL(reset_queue, dim3(1), dim3(1), 0, st, q);
    return hipGetLastError();
}

hipError_t launch_loglik(const ModelDev &M, const ChainArgs &A, int fp64, int BS, int grid, hipStream_t st)
{
    if (fp64) {
        if (BS == 64) return launch_t<double, 64>(M, A, grid, st);
        if (BS == 128) return launch_t<double, 128>(M, A, grid, st);
        if (BS == 256) return launch_t<double, 256>(M, A, grid, st);
    } else {
        if (BS == 64) return launch_t<float, 64>(M, A, grid, st);
        if (BS == 128) return launch_t<float, 128>(M, A, grid, st);
        if (BS == 256) return launch_t<float, 256>(M, A, grid, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace pa
