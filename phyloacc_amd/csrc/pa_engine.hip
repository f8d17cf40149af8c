// pa_engine.hip -- CDNA4 (gfx950) log-likelihood kernel: the element's
// log-likelihood at fixed ancestral states Z and rates (r1, r2), or at
// per-call branch lengths for an element with its own tree
// (BPP_C::log_f_Xz, src/PhyloAcc-ST/bpp_c2.cpp:392-425).  The MCMC chains
// are in pa_chain.hip.
//
// One workgroup owns one element at a time; workgroups pull elements from a
// device work queue (persistent kernel).  Within a workgroup:
//   * each branch's 4x4 transition matrix P(t * rate of its class) is built
//     in FP64 by one thread and stored in LDS in the partials' precision;
//   * threads stride over site patterns; a thread keeps the same patterns
//     through the whole post-order pass, so no barrier is needed between
//     tree nodes;
//   * conditional likelihoods are pattern-major in the workgroup's HBM slot,
//     linear-space with a power-of-two rescale per node; the cumulative
//     exponent rides along as an int; GAPT marks all-gap subtrees (the
//     reference's Tg >= num_base, src/PhyloAcc-ST/bpp_c.hpp:417-428);
//   * a cherry (a non-root internal node whose children are both leaves) is
//     never stored: its parent recomputes its scaled partial from the two
//     leaf codes with the same operations, so the first tree level costs no
//     pass of its own;
//   * the root sum over patterns is FP64 with a fixed reduction order.
// No MFMA: a K=4 matvec per pattern is not a dense contraction.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pa_common.h"
#include "pa_kernels.h"
#include "pa_device.h"
#include "pa_eig.h"
#include "pa_rng.h"

#include <mutex>

namespace pa {

// Scratch of the allocation-free likelihood entry point: the work counter in
// its first 256 bytes, then one slot per workgroup.  Every GPU the library
// runs on holds its own copy until the process ends.  Host code must never
// use the array's name as an address (on the host it is a stand-in symbol);
// it uses the device address from noalloc_scratch().
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

hipError_t noalloc_scratch(int device, char **ptr)
{
    static std::mutex mu;
    static char *addr[64] = {nullptr};
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        // the caller has made `device` current
        void *p = nullptr;
        const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess) return err;
        addr[device] = (char *)p;
    }
    *ptr = addr[device];
    return hipSuccess;
}

template <typename R>
struct Ctx {
    // element
    int S, N, P, Ppad, root, n_nodes, n_up;
    int h0;                    // first tree level the pass visits
    int path;                  // 1: the levels are one path (up[k]'s child on the path is up[k - 1])
    const GBL uint8_t *codes;
    const GBL int32_t *wts;
    int pre0, pre1;            // site-prefix root sums (pre0 < 0: the pattern sum)
    double sum1;               // the second prefix's sum
    // workgroup slot in HBM
    GBL V4<R> *parts;      // [I][Ppad]
    GBL int32_t *exps;     // [I][Ppad]
    // LDS arrays
    LDS int16_t *ch0, *ch1, *nodes, *up, *upoff;
    LDS uint8_t *flags;
    LDS R *PT;             // [N][16], PT[i*4+j] = P_{j->i}
    double red[16];
    double pi[4];          // base frequencies of the root sum (the model's, or the element's own)
    double Pc[16];         // a candidate's transition matrix, Pc[a*4+b] = P_{a->b} (gene-tree moves)
    const GBL int32_t *site;   // the element's site -> pattern map
    int work;
};

template <typename R>
using CP = LDS Ctx<R> *;
using MP = LDS const ModelDev *;
using AP = LDS const LoglikArgs *;

__host__ __device__ inline size_t al(size_t x, size_t a) { return (x + a - 1) & ~(a - 1); }

// Dynamic-LDS layout: returns the byte size; BIND also points the context's
// array fields into the block.  The context may sit at LDS address 0, which
// compares equal to a null pointer, so binding is selected at compile time and
// never by testing the context pointer.
template <typename R, bool BIND>
__host__ __device__ size_t smem_arrays(int N, LDS char *base, CP<R> c)
{
    size_t o = 0;
#define PA_TAKE(field, T, count)                                                   \
    do {                                                                           \
        o = al(o, 32);  /* rows of PT are read as 4-vectors */                      \
        if constexpr (BIND) c->field = (LDS T *)(base + o);                        \
        o += sizeof(T) * (size_t)(count);                                          \
    } while (0)
    PA_TAKE(PT, R, 16 * N);
    PA_TAKE(ch0, int16_t, N);
    PA_TAKE(ch1, int16_t, N);
    PA_TAKE(nodes, int16_t, N);
    PA_TAKE(up, int16_t, N);
    PA_TAKE(upoff, int16_t, N + 1);
    PA_TAKE(flags, uint8_t, N);
#undef PA_TAKE
    return al(o, 16);
}

size_t loglik_dyn_smem_bytes(int N, int fp64)
{
    return fp64 ? smem_arrays<double, false>(N, nullptr, nullptr) : smem_arrays<float, false>(N, nullptr, nullptr);
}

size_t loglik_smem_bytes(int N, int fp64)
{
    return (fp64 ? smem_arrays<double, false>(N, nullptr, nullptr) : smem_arrays<float, false>(N, nullptr, nullptr)) +
           (fp64 ? sizeof(Ctx<double>) : sizeof(Ctx<float>)) + sizeof(ModelDev) + sizeof(LoglikArgs);
}

// Byte offset of the exponents in a slot of I node slots, and the slot size.
__host__ __device__ inline size_t slot_exps_off(int I, int Ppad, int fp64)
{
    return al((size_t)I * Ppad * 4 * (fp64 ? 8 : 4), 256);
}

size_t loglik_slot_bytes(int N, int S, int Pcap, int fp64)
{
    return al(slot_exps_off(N - S, Pcap, fp64) + (size_t)(N - S) * Pcap * 4, 256);
}

size_t loglik_resident_bytes(int N, int S, int Ppad, int fp64)
{
    return al(slot_exps_off(2 * (N - S), Ppad, fp64) + (size_t)2 * (N - S) * Ppad * 4, 256);
}

// Node x's slot (x >= S) under its flags f: its first, or with NF_ALT its second.
__device__ __forceinline__ size_t node_slot(int x, int f, int S, int N)
{
    return (size_t)(x - S + ((f & NF_ALT) ? N - S : 0));
}

// ------------------------------------------------------------------ helpers

// Transition matrix of branch s into PT: P(d * rate) for length d > 1e-10,
// else the identity (BPP_C::initMats, src/PhyloAcc-ST/bpp_c.cpp:187-228).
// P(t) = V diag(exp(lam t)) Vi, clamped at 0, built in FP64, from the
// model's eigensystem or a gBGC class's (lam, V, Vi in LDS or global memory).
template <typename R, typename PE>
__device__ void branch_matrix_from(PE lam, PE V, PE Vi, CP<R> c, int s, double d, double rate)
{
    double P[16];
    if (!(d <= 1e-10)) {          // a NaN length gives a NaN matrix, not the identity
        const double t = d * rate;
        double e[4];
#pragma unroll
        for (int k = 0; k < 4; k++) e[k] = exp(lam[k] * t);
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                double v = 0;
#pragma unroll
                for (int k = 0; k < 4; k++) v += V[j * 4 + k] * e[k] * Vi[k * 4 + i];
                P[j * 4 + i] = v < 0 ? 0.0 : v;
            }
    } else {
#pragma unroll
        for (int q = 0; q < 16; q++) P[q] = (q % 5 == 0) ? 1.0 : 0.0;
    }
    LDS R *dst = c->PT + 16 * s;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) dst[i * 4 + j] = (R)P[j * 4 + i];
}

template <typename R>
__device__ void branch_matrix(MP M, CP<R> c, int s, double d, double rate)
{
    branch_matrix_from<R>(M->lam, M->V, M->Vi, c, s, d, rate);
}

// gBGC classes (BPP_C::init_cache_TM, src/PhyloAcc-ST-GBGC/SRC/bpp_c.cpp:
// 1563-1621, with the rate matrices of Gibbs, bpp_c.cpp:191-197): one thread
// per (work position, class).
struct GbgcEigArgs {
    double Q0[16];
    const double *sel;
    int n;
    double *eig;
};

__global__ void gbgc_eig_kernel(GbgcEigArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n * GB_CLASSES) return;
    const int w = t / GB_CLASSES, k = t % GB_CLASSES;
    const double *sel = a.sel + (size_t)w * 3;
    const double s = k % 3 == 0 ? 0.0 : (k % 3 == 1 ? sel[0] : sel[1]);
    const double b = k >= 3 ? sel[2] : 0.0;
    double Q[16], lam[4], V[16], Vi[16];
    gbgc_rate(a.Q0, s, b, Q);
    const int st = reversible_eig4(Q, lam, V, Vi);
    double *out = a.eig + (size_t)t * GB_EIG_WORDS;
    for (int i = 0; i < 4; i++) out[i] = st == EIG_OK ? lam[i] : NAN;
    for (int i = 0; i < 16; i++) {
        out[4 + i] = V[i];
        out[20 + i] = Vi[i];
    }
}

// Per-element strand-symmetric base frequencies (BPP_C::sample_pi and
// getQmat, src/PhyloAcc-GT/bpp_c.cpp:1545-1548,160-185): one thread per work
// position.
struct GtEigArgs {
    double inst[6];
    const double *piA;
    int n;
    double *eig, *pi;
};

__global__ void gt_eig_kernel(GtEigArgs a)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.n) return;
    double Q[16], pi[4], lam[4], V[16], Vi[16];
    gt_strand_q(a.inst, a.piA[w], Q, pi);
    const int st = reversible_eig4(Q, lam, V, Vi);
    double *out = a.eig + (size_t)w * GB_EIG_WORDS;
    for (int i = 0; i < 4; i++) out[i] = st == EIG_OK ? lam[i] : NAN;
    for (int i = 0; i < 16; i++) {
        out[4 + i] = V[i];
        out[20 + i] = Vi[i];
    }
    for (int i = 0; i < 4; i++) a.pi[(size_t)w * 4 + i] = pi[i];
}

hipError_t launch_gt_eig(const double *inst, const double *piA, int n, double *eig, double *pi, hipStream_t st)
{
    GtEigArgs a;
    for (int i = 0; i < 6; i++) a.inst[i] = inst[i];
    a.piA = piA;
    a.n = n;
    a.eig = eig;
    a.pi = pi;
    hipLaunchKernelGGL(gt_eig_kernel, dim3((n + 127) / 128), dim3(128), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_gbgc_eig(const double *Q0, const double *sel, int n, double *eig, hipStream_t st)
{
    GbgcEigArgs a;
    for (int i = 0; i < 16; i++) a.Q0[i] = Q0[i];
    a.sel = sel;
    a.n = n;
    a.eig = eig;
    const int threads = n * GB_CLASSES;
    hipLaunchKernelGGL(gbgc_eig_kernel, dim3((threads + 127) / 128), dim3(128), 0, st, a);
    return hipGetLastError();
}

// ----------------------------------------------------------------- pruning

// Patterns handled per thread per round: all their loads are issued before
// the first is used, so a round costs one memory latency.
template <typename R>
struct Unroll { static constexpr int U = sizeof(R) == 4 ? 4 : 2; };

// The root sum of a pass (log_f_Xz, src/PhyloAcc-ST/bpp_c2.cpp:392-425):
// log sum_j pi_j L_j + exponent * ln 2 per pattern (or per site over the two
// prefixes), FP64, identical on every thread.
template <typename R, int BS>
__device__ __forceinline__ double root_sum(CP<R> c)
{
    const int tid = threadIdx.x;
    const int P = uni(c->P), Ppad = uni(c->Ppad), S = uni(c->S), N = uni(c->N);
    GBL V4<R> *parts = uni_ptr(c->parts);
    GBL int32_t *exps = uni_ptr(c->exps);
    constexpr int U = Unroll<R>::U;
    // root: log sum_j pi_j L_j + exponent * ln 2; all-gap columns contribute 0
    const int r = c->root;
    const int pre0 = uni(c->pre0);
    double acc = 0, acc1 = 0;
    if (pre0 >= 0 && r >= S) {
        // two site prefixes: each site reads its pattern's root partial,
        // which another wave may have written, so the pass must finish first
        __syncthreads();
        const int pre1 = uni(c->pre1);
        const GBL int32_t *site = uni_ptr(c->site);
        const size_t ro = node_slot(r, uni(c->flags[r]), S, N) * Ppad;
        const GBL V4<R> *rp = parts + ro;
        const GBL int32_t *re = exps + ro;
        const double p0 = c->pi[0], p1 = c->pi[1], p2 = c->pi[2], p3 = c->pi[3];
        for (int g = tid; g < pre0; g += BS) {
            const int p = site[g];
            const int ex = re[p];
            if (ex == GAPT) continue;
            const V4<R> av = ld4<R>(rp + p);
            const double L = p0 * (double)av.x + p1 * (double)av.y + p2 * (double)av.z + p3 * (double)av.w;
            const double t = log(L) + (double)ex * 0.69314718055994530942;
            acc += t;
            if (g < pre1) acc1 += t;
        }
        const double s1 = block_sum(acc1, c->red);
        if (threadIdx.x == 0) c->sum1 = s1;
        return block_sum(acc, c->red);
    }
    if (r >= S) {
        const size_t ro = node_slot(r, uni(c->flags[r]), S, N) * Ppad;
        const GBL V4<R> *rp = parts + ro;
        const GBL int32_t *re = exps + ro;
        const double p0 = c->pi[0], p1 = c->pi[1], p2 = c->pi[2], p3 = c->pi[3];
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
                const double t = log(L) + (double)ex[u] * 0.69314718055994530942;
                acc += (double)wt[u] * t;
            }
        }
    }
    return block_sum(acc, c->red);
}

// Node x's input at pattern p as the pruning pass loads it: a leaf's code, a
// cherry's two leaf codes packed, or a stored node's partial and exponent.
template <typename R>
__device__ __forceinline__ ChildIn<R> raw_in(CP<R> c, const GBL uint8_t *codes, const GBL V4<R> *parts,
                                             const GBL int32_t *exps, int x, int f, int S, int N, int Ppad, int p)
{
    ChildIn<R> in;
    in.v.x = in.v.y = in.v.z = in.v.w = (R)0;
    if (x < S) {
        in.x = codes[(size_t)x * Ppad + p];
    } else if (f & NF_CHERRY) {
        in.x = codes[(size_t)c->ch0[x] * Ppad + p] | (codes[(size_t)c->ch1[x] * Ppad + p] << 8);
    } else {
        const size_t o = node_slot(x, f, S, N) * Ppad;
        in.x = exps[o + p];
        if (!(f & NF_MISSING)) in.v = ld4<R>(parts + o + p);
    }
    return in;
}

// A cherry's packed codes into its scaled partial (cherry_in); other inputs as they are.
template <typename R>
__device__ __forceinline__ ChildIn<R> cooked_in(CP<R> c, int x, int f, ChildIn<R> in)
{
    if (!(f & NF_CHERRY)) return in;
    const int yA = c->ch0[x], yB = c->ch1[x];
    return cherry_in<R>(c->PT + 16 * yA, c->PT + 16 * yB, c->flags[yA] & NF_MISSING, c->flags[yB] & NF_MISSING,
                        f & NF_MISSING, in.x & 255, in.x >> 8);
}

// Node s at pattern p from its children's inputs (child 0 first, as the
// pruning pass applies them), stored into s's slot; returns what a later
// load of s at p would read.
template <typename R>
__device__ __forceinline__ ChildIn<R> node_at(CP<R> c, GBL V4<R> *parts, GBL int32_t *exps, int s, int x0, int x1,
                                              const ChildIn<R> &in0, const ChildIn<R> &in1, int S, int N, int Ppad,
                                              int p)
{
    const int fs = c->flags[s], f0 = c->flags[x0], f1 = c->flags[x1];
    const size_t o = node_slot(s, fs, S, N) * Ppad;
    R a0 = 1, a1 = 1, a2 = 1, a3 = 1;
    int e = 0;
    bool g0, g1;
    child_apply<R>(x0 < S, f0 & NF_MISSING, c->PT + 16 * x0, in0, a0, a1, a2, a3, e, g0);
    child_apply<R>(x1 < S, f1 & NF_MISSING, c->PT + 16 * x1, in1, a0, a1, a2, a3, e, g1);
    ChildIn<R> out;
    out.v.x = out.v.y = out.v.z = out.v.w = (R)0;
    if (g0 && g1) {
        out.x = GAPT;
    } else if (fs & NF_MISSING) {
        out.x = 0;
    } else {
        const R mx = fmax(fmax(a0, a1), fmax(a2, a3));
        int kx = 0;
        frexp(mx, &kx);
        a0 = ldexp(a0, -kx); a1 = ldexp(a1, -kx); a2 = ldexp(a2, -kx); a3 = ldexp(a3, -kx);
        st4<R>(parts + o + p, a0, a1, a2, a3);
        out.v.x = a0; out.v.y = a1; out.v.z = a2; out.v.w = a3;
        out.x = e + kx;
    }
    exps[o + p] = out.x;
    return out;
}

// The levels of a path pass (the height moves' resident partials): level k
// is the one node up[k], and its child on the path is up[k - 1].  A thread
// takes each of its patterns through every level: the child on the path
// comes from the previous level in registers, and the loads of the next
// levels' other children are in flight PD levels ahead, so the levels do not
// each wait for a memory round trip.  The values are the level-by-level
// pass's, node by node and pattern by pattern.
template <typename R, int BS>
__device__ __noinline__ double prune_path(CP<R> c, unsigned long long *bytes)
{
    constexpr int PD = sizeof(R) == 4 ? 8 : 4;
    const int tid = threadIdx.x;
    const int P = uni(c->P), Ppad = uni(c->Ppad), S = uni(c->S), N = uni(c->N), K = uni(c->n_up);
    GBL V4<R> *parts = uni_ptr(c->parts);
    GBL int32_t *exps = uni_ptr(c->exps);
    const GBL uint8_t *codes = uni_ptr(c->codes);
    if (bytes)
        for (int q = 0; q < K; q++) {
            // as a level-by-level pass counts its nodes
            const int s = uni(c->up[q]);
            unsigned long long per = 4 * sizeof(R) + 4;
            for (int j = 0; j < 2; j++) {
                const int x = uni(j ? c->ch1[s] : c->ch0[s]);
                per += x < S ? 1 : ((uni(c->flags[x]) & NF_CHERRY) ? 2 : 4 * sizeof(R) + 4);
            }
            *bytes += (unsigned long long)P * per;
        }
    for (int p = tid; p < P; p += BS) {
        // level 0: both children are read
        const int s0 = uni(c->up[0]);
        const int a0 = uni(c->ch0[s0]), a1 = uni(c->ch1[s0]);
        const int fa0 = uni(c->flags[a0]), fa1 = uni(c->flags[a1]);
        ChildIn<R> i0 = raw_in<R>(c, codes, parts, exps, a0, fa0, S, N, Ppad, p);
        ChildIn<R> i1 = raw_in<R>(c, codes, parts, exps, a1, fa1, S, N, Ppad, p);
        // the other children of levels 1 .. PD, in flight
        ChildIn<R> ring[PD];
#pragma unroll
        for (int j = 0; j < PD; j++) {
            const int h = 1 + j;
            if (h < K) {
                const int s = uni(c->up[h]), prev = uni(c->up[h - 1]);
                const int x0 = uni(c->ch0[s]), x1 = uni(c->ch1[s]);
                const int xo = x0 == prev ? x1 : x0;
                ring[j] = raw_in<R>(c, codes, parts, exps, xo, uni(c->flags[xo]), S, N, Ppad, p);
            }
        }
        ChildIn<R> carry = node_at<R>(c, parts, exps, s0, a0, a1, cooked_in<R>(c, a0, fa0, i0),
                                      cooked_in<R>(c, a1, fa1, i1), S, N, Ppad, p);
        for (int h0 = 1; h0 < K; h0 += PD) {
#pragma unroll
            for (int j = 0; j < PD; j++) {
                const int h = h0 + j;
                if (h >= K) break;
                const int s = uni(c->up[h]), prev = uni(c->up[h - 1]);
                const int x0 = uni(c->ch0[s]), x1 = uni(c->ch1[s]);
                const bool on0 = x0 == prev;
                const int xo = on0 ? x1 : x0;
                const ChildIn<R> other = cooked_in<R>(c, xo, uni(c->flags[xo]), ring[j]);
                if (h + PD < K) {
                    const int sn = uni(c->up[h + PD]), pn = uni(c->up[h + PD - 1]);
                    const int y0 = uni(c->ch0[sn]), y1 = uni(c->ch1[sn]);
                    const int yo = y0 == pn ? y1 : y0;
                    ring[j] = raw_in<R>(c, codes, parts, exps, yo, uni(c->flags[yo]), S, N, Ppad, p);
                }
                carry = on0 ? node_at<R>(c, parts, exps, s, x0, x1, carry, other, S, N, Ppad, p)
                            : node_at<R>(c, parts, exps, s, x0, x1, other, carry, S, N, Ppad, p);
            }
        }
    }
    return root_sum<R, BS>(c);
}

// The pruning pass over every internal node of the element's tree
// (BPP_C::Update_Tg(tosample=false), src/PhyloAcc-ST/bpp_c.cpp:644-671) and
// the root sum (log_f_Xz, src/PhyloAcc-ST/bpp_c2.cpp:392-425).  Missing
// nodes only record their GAPT flag.  Returns the element's log-likelihood
// (FP64), identical on every thread.
//
// A thread owns the patterns p = tid (mod BS) of every node, so it only ever
// reads partials it wrote itself and the pass needs no barrier.  Nodes of one
// tree level are independent: a round takes U (node, pattern block) items of
// the same level and issues all their loads before the first is used, so a
// round costs one memory latency whether the level is wide (many nodes) or
// the element is long (many pattern blocks).
template <typename R, int BS>
__device__ __noinline__ double prune(MP M, CP<R> c, unsigned long long *bytes)
{
    const int tid = threadIdx.x;
    const int P = uni(c->P), Ppad = uni(c->Ppad), S = uni(c->S), N = uni(c->N);
    GBL V4<R> *parts = uni_ptr(c->parts);
    GBL int32_t *exps = uni_ptr(c->exps);
    const GBL uint8_t *codes = uni_ptr(c->codes);
    const int Pb = uni((P + BS - 1) / BS);
    constexpr int U = Unroll<R>::U;
    const int n_up = uni(c->n_up);
    // every node of the first level (height 1) is a cherry, unless the tree is
    // one cherry: its root
    for (int h = uni(c->h0); h < n_up; h++) {
        // level h: up[upoff[h-1] .. upoff[h]) (from 0 for the first level)
        const int a = uni(h == 0 ? 0 : c->upoff[h - 1]), b = uni(c->upoff[h]);
        const int items = (b - a) * Pb;
        if (bytes) {
            // write partial + exponent; read each child's partial + exponent,
            // its leaf code, or a cherry's two leaf codes
            for (int q = a; q < b; q++) {
                const int s = uni(c->up[q]);
                unsigned long long per = 4 * sizeof(R) + 4;
                for (int j = 0; j < 2; j++) {
                    const int x = uni(j ? c->ch1[s] : c->ch0[s]);
                    per += x < S ? 1 : ((uni(c->flags[x]) & NF_CHERRY) ? 2 : 4 * sizeof(R) + 4);
                }
                *bytes += (unsigned long long)P * per;
            }
        }
        for (int i0 = 0; i0 < items; i0 += U) {
            ChildIn<R> in0[U], in1[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u;
                if (i >= items) break;
                const int q = i / Pb, p = (i - q * Pb) * BS + tid;
                const int s = uni(c->up[a + q]);
                const int x0 = uni(c->ch0[s]), x1 = uni(c->ch1[s]);
                const bool leaf0 = x0 < S, leaf1 = x1 < S;
                const int f0 = uni(c->flags[x0]), f1 = uni(c->flags[x1]);
                const bool miss0 = f0 & NF_MISSING, miss1 = f1 & NF_MISSING;
                const size_t o0 = leaf0 ? 0 : node_slot(x0, f0, S, N) * Ppad;
                const size_t o1 = leaf1 ? 0 : node_slot(x1, f1, S, N) * Ppad;
                if (p < P) {
                    // a cherry child: its two leaf codes, packed
                    if (f0 & NF_CHERRY)
                        in0[u].x = codes[(size_t)uni(c->ch0[x0]) * Ppad + p] | (codes[(size_t)uni(c->ch1[x0]) * Ppad + p] << 8);
                    else
                        in0[u] = child_load<R>(leaf0, miss0, codes + (size_t)x0 * Ppad, parts + o0, exps + o0, p);
                    if (f1 & NF_CHERRY)
                        in1[u].x = codes[(size_t)uni(c->ch0[x1]) * Ppad + p] | (codes[(size_t)uni(c->ch1[x1]) * Ppad + p] << 8);
                    else
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
                if (p >= P) continue;
                const int s = uni(c->up[a + q]);
                const int x0 = uni(c->ch0[s]), x1 = uni(c->ch1[s]);
                const bool leaf0 = x0 < S, leaf1 = x1 < S;
                const int fs = uni(c->flags[s]);
                const bool miss_s = fs & NF_MISSING;
                const int f0 = uni(c->flags[x0]), f1 = uni(c->flags[x1]);
                const bool miss0 = f0 & NF_MISSING, miss1 = f1 & NF_MISSING;
                if (f0 & NF_CHERRY) {
                    const int yA = uni(c->ch0[x0]), yB = uni(c->ch1[x0]);
                    in0[u] = cherry_in<R>(c->PT + 16 * yA, c->PT + 16 * yB, uni(c->flags[yA]) & NF_MISSING,
                                          uni(c->flags[yB]) & NF_MISSING, miss0, in0[u].x & 255, in0[u].x >> 8);
                }
                if (f1 & NF_CHERRY) {
                    const int yA = uni(c->ch0[x1]), yB = uni(c->ch1[x1]);
                    in1[u] = cherry_in<R>(c->PT + 16 * yA, c->PT + 16 * yB, uni(c->flags[yA]) & NF_MISSING,
                                          uni(c->flags[yB]) & NF_MISSING, miss1, in1[u].x & 255, in1[u].x >> 8);
                }
                GBL V4<R> *dstp = parts + node_slot(s, fs, S, N) * Ppad;
                GBL int32_t *dste = exps + node_slot(s, fs, S, N) * Ppad;
                R a0 = 1, a1 = 1, a2 = 1, a3 = 1;
                int e = 0;
                bool g0, g1;
                child_apply<R>(leaf0, miss0, c->PT + 16 * x0, in0[u], a0, a1, a2, a3, e, g0);
                child_apply<R>(leaf1, miss1, c->PT + 16 * x1, in1[u], a0, a1, a2, a3, e, g1);
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
    return root_sum<R, BS>(c);
}

// ------------------------------------------- gene-tree moves (PhyloAcc-GT)

// Node v's scaled partial and exponent at pattern p (GAPT: no base below);
// a cherry's is recomputed from its leaf codes, as the pruning pass does.
template <typename R>
__device__ inline ChildIn<R> node_in(CP<R> c, int v, int p)
{
    const int S = c->S, Ppad = c->Ppad;
    if (c->flags[v] & NF_CHERRY) {
        const int yA = c->ch0[v], yB = c->ch1[v];
        return cherry_in<R>(c->PT + 16 * yA, c->PT + 16 * yB, c->flags[yA] & NF_MISSING, c->flags[yB] & NF_MISSING,
                            c->flags[v] & NF_MISSING, c->codes[(size_t)yA * Ppad + p], c->codes[(size_t)yB * Ppad + p]);
    }
    ChildIn<R> in;
    const size_t o = node_slot(v, c->flags[v], S, c->N) * Ppad;
    in.x = c->exps[o + p];
    in.v = ld4<R>(c->parts + o + p);
    return in;
}

// A leaf code as a history base: 0-3, -1 for an ambiguity code, 4 for a gap
// or another character (src/PhyloAcc-GT/bpp_c.hpp:340-384).
__device__ inline int leaf_base(int code)
{
    return code < 4 ? code : (code == PA_GAP_CODE || code == PA_OTHER_CODE ? 4 : -1);
}

// GTree::Update_Tg with tosample = true (src/PhyloAcc-GT/genetree.cpp:
// 1416-1532) after the pruning pass: one thread per site draws the gene
// root's base from pi * L_root, then every internal gene node's from
// P(parent base -> .) * L over its whole branch (the law of the reference's
// species-by-species steps), pre-order, and writes every node's base.
template <typename R, int BS>
__device__ __noinline__ void gt_history(MP M, CP<R> c, AP A, int row)
{
    const int N = c->N, S = c->S, Ppad = c->Ppad, root = c->root, n_nodes = c->n_nodes;
    const int len = A->hist_len[row], stride = A->hist_ns[row];
    GBL int8_t *out = (GBL int8_t *)A->hist + A->hist_off[row];
    const GBL int32_t *site = c->site;
    const Key key{M->key0, M->key1};
    const uint32_t c2 = (uint32_t)A->hist_sweep, c3 = tagword(0, TAG_GT_HIST);
    for (int g = threadIdx.x; g < len; g += BS) {
        const int p = site[g];
        for (int v = 0; v < S; v++) out[(size_t)v * stride + g] = (int8_t)leaf_base(c->codes[(size_t)v * Ppad + p]);
        {
            const ChildIn<R> in = node_in<R>(c, root, p);
            const bool none = in.x == GAPT || (c->flags[root] & NF_MISSING);
            double w[4] = {c->pi[0], c->pi[1], c->pi[2], c->pi[3]};
            if (!none) { w[0] *= (double)in.v.x; w[1] *= (double)in.v.y; w[2] *= (double)in.v.z; w[3] *= (double)in.v.w; }
            out[(size_t)root * stride + g] = (int8_t)categorical<4>(w, slot_uniform(key, (uint32_t)g, c2, c3, (uint32_t)root));
        }
        for (int q = n_nodes - 1; q >= 0; q--) {
            const int u = c->nodes[q];
            if (u < S) continue;
            const int xu = out[(size_t)u * stride + g];
            for (int k = 0; k < 2; k++) {
                const int v = k ? c->ch1[u] : c->ch0[u];
                if (v < S) continue;
                int res = 4;
                if (!(c->flags[v] & NF_MISSING) && xu >= 0 && xu < 4) {
                    const ChildIn<R> in = node_in<R>(c, v, p);
                    if (in.x != GAPT) {
                        const LDS R *PT = c->PT + 16 * v;      // PT[i*4+j] = P_{j->i}
                        double w[4] = {(double)PT[0 * 4 + xu] * (double)in.v.x, (double)PT[1 * 4 + xu] * (double)in.v.y,
                                       (double)PT[2 * 4 + xu] * (double)in.v.z, (double)PT[3 * 4 + xu] * (double)in.v.w};
                        res = categorical<4>(w, slot_uniform(key, (uint32_t)g, c2, c3, (uint32_t)v));
                    }
                }
                out[(size_t)v * stride + g] = (int8_t)res;
            }
        }
    }
}

// P(d) = V e^{lam d} Vi in FP64 into Pc[a*4+b] = P_{a->b}, clamped at 0; the
// identity below 1e-8 (BPP::getlogTM_len, src/PhyloAcc-GT/bpp.cpp:238-252).
template <typename PE>
__device__ inline void full_transition(PE lam, PE V, PE Vi, double d, LDS double *Pc)
{
    for (int a = 0; a < 4; a++)
        for (int j = 0; j < 4; j++) {
            double v = 0;
            for (int k = 0; k < 4; k++) v += V[a * 4 + k] * exp(lam[k] * d) * Vi[k * 4 + j];
            Pc[a * 4 + j] = d < 1e-8 ? (a == j ? 1.0 : 0.0) : (v < 0 ? 0.0 : v);
        }
}

// Candidate scores of GTree::Sample_tree2 (src/PhyloAcc-GT/genetree.cpp:
// 2190-2380) after the pruning pass of the current tree: for each candidate
// new grandparent of branch b at substitution distance d, the sum over sites
// where b holds a base of: kind 0, log (P(d) L_b)[x_t] at the candidate's
// sampled base x_t (the pi-average of P(d) L_b when it has none); kind 1, the
// pi-average; kind 2, log sum_i pi_i P(d)_{x_b -> i} at b's own base.
template <typename R, int BS>
__device__ __noinline__ void gt_candidates(MP M, CP<R> c, AP A, int w, int row)
{
    const int S = c->S, Ppad = c->Ppad, stride = A->hist_ns[row], cm = A->cand_max;
    const GBL int8_t *tg = (const GBL int8_t *)A->hist + A->hist_off[row];
    const int b = A->cand_branch[row], nc = A->cand_n[row], len = A->cand_sites[row];
    const GBL int32_t *site = c->site;
    const double *E = A->el_eig ? A->el_eig + (size_t)w * GB_EIG_WORDS : nullptr;
    const double lp[4] = {log(c->pi[0]), log(c->pi[1]), log(c->pi[2]), log(c->pi[3])};
    for (int q = 0; q < nc; q++) {
        const int t = A->cand_node[(size_t)row * cm + q], kind = A->cand_kind[(size_t)row * cm + q];
        if (threadIdx.x == 0) {
            const double d = A->cand_len[(size_t)row * cm + q];
            if (E) full_transition(E, E + 4, E + 20, d, c->Pc);
            else full_transition(M->lam, M->V, M->Vi, d, c->Pc);
        }
        __syncthreads();
        double acc = 0;
        for (int g = threadIdx.x; g < len; g += BS) {
            const int xb = tg[(size_t)b * stride + g];
            if (xb >= 4) continue;
            if (kind == 2) {
                if (xb < 0) continue;
                double sum = 0;
                for (int i = 0; i < 4; i++) sum += c->pi[i] * c->Pc[xb * 4 + i];
                acc += log(sum);
                continue;
            }
            const int p = site[g];
            double L[4];
            int e = 0;
            if (b < S) {
                const int code = c->codes[(size_t)b * Ppad + p];
                for (int j = 0; j < 4; j++) L[j] = 0;
                if (code < 4) L[code] = 1;
                else {
                    const int b1 = (code == 6 || code == 9 || code == 11) ? 0 : (code == 7 || code == 10) ? 1 : 2;
                    const int b2 = (code == 6 || code == 10) ? 2 : (code == 9) ? 1 : 3;
                    L[b1] = 1; L[b2] = 1;
                }
            } else {
                const ChildIn<R> in = node_in<R>(c, b, p);
                L[0] = in.v.x; L[1] = in.v.y; L[2] = in.v.z; L[3] = in.v.w;
                e = in.x;
            }
            double Mv[4];
            for (int a = 0; a < 4; a++)
                Mv[a] = log(c->Pc[a * 4 + 0] * L[0] + c->Pc[a * 4 + 1] * L[1] + c->Pc[a * 4 + 2] * L[2] +
                            c->Pc[a * 4 + 3] * L[3]) + (double)e * 0.69314718055994530942;
            const int xt = kind == 0 ? tg[(size_t)t * stride + g] : 4;
            if (kind == 0 && !(c->flags[t] & NF_MISSING) && xt >= 0 && xt < 4) {
                acc += Mv[xt];
            } else {
                double mx = -INFINITY;
                for (int a = 0; a < 4; a++) mx = fmax(mx, Mv[a] + lp[a]);
                double sum = 0;
                for (int a = 0; a < 4; a++) sum += exp(Mv[a] + lp[a] - mx);
                acc += mx == -INFINITY ? -INFINITY : log(sum) + mx;
            }
        }
        const double tot = block_sum(acc, c->red);
        if (threadIdx.x == 0) A->cand_score[(size_t)row * cm + q] = tot;
        __syncthreads();
    }
}

// ---------------------------------------------------------- element setup

// Binds work position w: the element's node table into LDS, its slot, and
// the transition matrix of every branch.  A branch takes the rate of its
// class: Z = 0 (neutral) rate 1, Z = 2 (accelerated) r2, any other Z (1,
// conserved) r1; an own-tree element takes its lengths from len_in.
template <typename R, int BS>
__device__ __noinline__ void bind_element(MP M, CP<R> c, AP A, int w)
{
    const int tid = threadIdx.x, N = M->N;
    const ElemDev E = A->elems[A->work[w]];
    const int row = A->len_row ? A->len_row[w] : w;
    const int32_t *hdr = A->hdr_call ? A->hdr_call + (size_t)row * 4 : nullptr;
    if (tid == 0) {
        c->S = M->S; c->N = N;
        c->P = E.P; c->Ppad = E.Ppad; c->root = hdr ? hdr[0] : E.root;
        c->site = A->site ? (const GBL int32_t *)(A->site + E.off_site) : nullptr;
        const double *pw = A->el_pi ? A->el_pi + (size_t)w * 4 : nullptr;
        for (int j = 0; j < 4; j++) c->pi[j] = pw ? pw[j] : M->pi[j];
        c->n_nodes = hdr ? hdr[1] : E.n_nodes;
        c->n_up = hdr ? hdr[2] : E.n_up;
        c->codes = (const GBL uint8_t *)(A->codes + E.off_codes);
        c->wts = (const GBL int32_t *)(A->wts + E.off_w);
        c->pre0 = A->pre_len ? A->pre_len[2 * row] : -1;
        c->pre1 = A->pre_len ? A->pre_len[2 * row + 1] : -1;
        const int64_t bo = A->big_off ? A->big_off[w] : -1;
        char *base = bo >= 0 ? A->big + bo : A->scratch + (int64_t)blockIdx.x * A->slot_bytes;
        c->parts = (GBL V4<R> *)base;
        c->exps = (GBL int32_t *)(base + slot_exps_off((A->alt ? 2 : 1) * (N - M->S), E.Ppad, sizeof(R) == 8));
    }
    const int32_t *tab = A->tab_call ? A->tab_call + (size_t)row * tab_words(N) : A->tab + E.off_tab;
    for (int i = tid; i < N; i += BS) {
        c->ch0[i] = (int16_t)tab[tab_ch0(N) + i];
        c->ch1[i] = (int16_t)tab[tab_ch1(N) + i];
        c->flags[i] = (uint8_t)tab[tab_flags(N) + i];
        c->nodes[i] = (int16_t)tab[tab_nodes(N) + i];
        c->up[i] = (int16_t)tab[tab_up(N) + i];
    }
    for (int i = tid; i <= N; i += BS) c->upoff[i] = (int16_t)tab[tab_upoff(N) + i];
    __syncthreads();
    // cherries (both children leaves; never the root) are recomputed where read
    if (c->n_up > 1)
        for (int q = tid; q < c->upoff[0]; q += BS) c->flags[c->up[q]] |= NF_CHERRY;
    __syncthreads();
    if (A->alt)
        for (int i = tid; i < N; i += BS)
            if (A->alt[(size_t)row * N + i]) c->flags[i] |= NF_ALT;
    __syncthreads();
    if (tid == 0) {
        c->h0 = c->n_up > 1 ? 1 : 0;
        c->path = 0;
        const int v = A->path_node ? A->path_node[row] : -1;
        if (v >= c->S && v < N) {
            // the pass visits the path from v to the root (cherries are
            // recomputed where read), one node per level, writing each into
            // the slot it does not hold now
            int k = 0;
            for (int x = v; x >= 0 && x < N && k < N; x = tab[tab_par(N) + x]) {
                c->flags[x] ^= NF_ALT;
                if (!(c->flags[x] & NF_CHERRY)) {
                    c->up[k] = (int16_t)x;
                    c->upoff[k] = (int16_t)(k + 1);
                    k++;
                }
                if (x == c->root) break;
            }
            c->n_up = k;
            c->h0 = 0;
            c->path = 1;
        }
    }
    const bool own = A->len_in != nullptr, gb = A->gb_eig != nullptr;
    const double r1 = own || gb ? 1.0 : A->rates_in[2 * w], r2 = own || gb ? 1.0 : A->rates_in[2 * w + 1];
    const double *dist = (const double *)(tab + tab_dist(N));
    for (int q = tid; q < c->n_nodes - 1; q += BS) {
        const int s = c->nodes[q];
        if (own) {
            const double len = A->len_in[(size_t)(A->len_row ? A->len_row[w] : w) * N + s];
            if (A->el_eig) {
                const double *E = A->el_eig + (size_t)w * GB_EIG_WORDS;
                branch_matrix_from<R>(E, E + 4, E + 20, c, s, len, 1.0);
            } else {
                branch_matrix<R>(M, c, s, len, 1.0);
            }
        } else if (gb) {
            const int z = A->Zin[(size_t)w * N + s];
            const int k = (z == 2 ? 2 : (z == 0 ? 0 : 1)) + (A->Bin[(size_t)w * N + s] != 0 ? 3 : 0);
            const double *E = A->gb_eig + ((size_t)w * GB_CLASSES + k) * GB_EIG_WORDS;
            branch_matrix_from<R>(E, E + 4, E + 20, c, s, dist[s], 1.0);
        } else {
            const int z = A->Zin[(size_t)w * N + s];
            branch_matrix<R>(M, c, s, dist[s], z == 2 ? r2 : (z == 0 ? 1.0 : r1));
        }
    }
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

// Declared for 3 waves per SIMD (at most 168 VGPRs): the pruning pass then
// keeps all its rounds' loads in registers, and only the element setup saves
// registers at its entry.  Measured on 120-leaf gene trees (20,000 elements,
// FP32): 874 K element log-likelihoods/s, against 761 K at the unbounded
// 228 VGPRs (2 waves) and 863 K at 4 waves, where the setup spills.
template <typename R, int BS>
__global__ void __launch_bounds__(BS, 3) loglik_kernel(ModelDev Mk, LoglikArgs Ak)
{
    extern __shared__ __align__(16) char smem_raw[];
    __shared__ ModelDev Ms;
    __shared__ LoglikArgs As;
    __shared__ Ctx<R> Cs;
    MP M = to_lds(Mk, (LDS ModelDev *)&Ms);
    AP A = to_lds(Ak, (LDS LoglikArgs *)&As);
    CP<R> c = (CP<R>)&Cs;
    __syncthreads();
    if (threadIdx.x == 0) smem_arrays<R, true>(M->N, (LDS char *)smem_raw, c);
    __syncthreads();
    for (;;) {
        if (threadIdx.x == 0) c->work = atomicAdd(A->queue, 1);
        __syncthreads();
        const int w = c->work;
        __syncthreads();
        if (w >= A->n_work) break;
        // an id outside the arenas, or a filtered element (no patterns), scores NaN
        const int id = A->work[w];
        if (id < 0 || id >= A->n_elems || A->elems[id].P <= 0) {
            if (threadIdx.x == 0) A->ll_out[w] = __builtin_nan("");
            __syncthreads();
            continue;
        }
        bind_element<R, BS>(M, c, A, w);
        unsigned long long bytes = 0;
        unsigned long long *bp = A->counters && threadIdx.x == 0 ? &bytes : nullptr;
        const double l = uni(c->path) ? prune_path<R, BS>(c, bp) : prune<R, BS>(M, c, bp);
        if (A->hist_draw || A->cand_score) {
            const int row = A->len_row ? A->len_row[w] : w;
            if (A->hist_draw) gt_history<R, BS>(M, c, A, row);
            if (A->cand_score) gt_candidates<R, BS>(M, c, A, w, row);
        }
        if (threadIdx.x == 0) {
            A->ll_out[w] = l;
            if (A->pre_len) A->ll_out2[w] = c->sum1;
            if (A->row_out0) {
                const int row = A->len_row ? A->len_row[w] : w;
                A->row_out0[row] = l;
                if (A->pre_len) A->row_out1[row] = c->sum1;
            }
            // root partial + exponent + pattern weight
            if (A->counters) atomicAdd(A->counters, bytes + (unsigned long long)c->P * (4 * sizeof(R) + 8));
        }
        __syncthreads();
    }
}

__global__ void reset_queue(int *q) { *q = 0; }

template <typename R, int BS>
static hipError_t launch_t(const ModelDev &M, const LoglikArgs &A, int grid, hipStream_t st)
{
    // dynamic LDS above 64 KB needs the per-kernel limit raised (large trees only)
    const size_t sm = smem_arrays<R, false>(M.N, nullptr, nullptr);
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

template <typename R, int BS>
static int blocks_t(int N)
{
    const size_t sm = smem_arrays<R, false>(N, nullptr, nullptr);
    if (sm > 65536 &&
        hipFuncSetAttribute((const void *)loglik_kernel<R, BS>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm) != hipSuccess)
        return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, loglik_kernel<R, BS>, BS, sm) != hipSuccess) return 0;
    return nb;
}

int loglik_blocks_per_cu(int fp64, int BS, int N)
{
    if (fp64) {
        if (BS == 64) return blocks_t<double, 64>(N);
        if (BS == 128) return blocks_t<double, 128>(N);
        if (BS == 256) return blocks_t<double, 256>(N);
    } else {
        if (BS == 64) return blocks_t<float, 64>(N);
        if (BS == 128) return blocks_t<float, 128>(N);
        if (BS == 256) return blocks_t<float, 256>(N);
    }
    return 0;
}

hipError_t launch_loglik(const ModelDev &M, const LoglikArgs &A, int fp64, int BS, int grid, hipStream_t st)
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
