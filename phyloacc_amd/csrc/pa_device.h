// pa_device.h -- device helpers shared by the chain kernel (pa_chain.hip)
// and the log-likelihood kernel (pa_engine.hip): 4-vectors, the per-pattern
// pruning steps (leaf rows, cherries recomputed from their leaf codes, child
// messages) and wave-uniform values.  One copy, so the two kernels apply the
// same operations in the same order and their FP32 results agree bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pa_common.h"
#include "pa_kernels.h"   // leaf codes PA_GAP_CODE, PA_OTHER_CODE

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

// Values every lane of the workgroup holds alike (node ids, flags, offsets
// read from LDS) are moved to scalar registers, which frees vector registers
// and makes the address arithmetic scalar.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
// Global pointers are 64-bit, LDS pointers 32-bit.
template <typename T>
__device__ __forceinline__ T *uni_ptr(T *p)
{
    const uintptr_t v = (uintptr_t)p;
    if constexpr (sizeof(T *) == 4) {
        return (T *)(uintptr_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    } else {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
        return (T *)(((uint64_t)hi << 32) | lo);
    }
}

// Deterministic workgroup sum: a fixed xor tree inside each wave, then the
// waves' sums in order (red: one double per wave, in LDS).
__device__ __forceinline__ double block_sum(double v, LDS double *red)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0;
    for (int i = 0; i < nw; i++) t += red[i];
    __syncthreads();
    return t;
}

// Node flag set when an element is bound (above the element table's flags):
// the node is a cherry (a non-root internal node whose children are both
// leaves), whose partial is recomputed where it is read.
constexpr uint8_t NF_CHERRY = 0x80;
// Node flag of resident partials: the node's partial lives in its second
// slot (at node + N - S) rather than its first.
constexpr uint8_t NF_ALT = 0x40;

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

// The scaled partial and exponent of a cherry at one pattern, from its leaf
// codes cA, cB and their branch matrices: the operations a pruning pass
// applies to any node, in the same order, so the values are those the pass
// would have stored.
template <typename R>
__device__ __forceinline__ ChildIn<R> cherry_in(const LDS R *srcA, const LDS R *srcB, bool missA, bool missB,
                                                bool miss_x, int cA, int cB)
{
    ChildIn<R> in;
    in.v.x = in.v.y = in.v.z = in.v.w = (R)0;
    R a0 = 1, a1 = 1, a2 = 1, a3 = 1;
    const bool gA = (cA == PA_GAP_CODE || cA == PA_OTHER_CODE);
    const bool gB = (cB == PA_GAP_CODE || cB == PA_OTHER_CODE);
    if (!missA && !gA) leaf_mul<R>(srcA, cA, a0, a1, a2, a3);
    if (!missB && !gB) leaf_mul<R>(srcB, cB, a0, a1, a2, a3);
    if (gA && gB) {
        in.x = GAPT;
    } else if (miss_x) {
        in.x = 0;
    } else {
        const R mx = fmax(fmax(a0, a1), fmax(a2, a3));
        int kx = 0;
        frexp(mx, &kx);
        in.v.x = ldexp(a0, -kx); in.v.y = ldexp(a1, -kx); in.v.z = ldexp(a2, -kx); in.v.w = ldexp(a3, -kx);
        in.x = kx;
    }
    return in;
}

// Multiplies one child's message P * L_child into (a0..a3).  A leaf reads the
// row of its base from the child's LDS matrix PT; an internal child reads the
// whole matrix as four 4-vectors.  gap reports the child's GAPT state
// (the reference's Tg >= num_base).
template <typename R>
__device__ __forceinline__ void child_apply(bool leaf, bool miss, const LDS R *PT, const ChildIn<R> &in,
                                            R &a0, R &a1, R &a2, R &a3, int &e, bool &gap)
{
    if (leaf) {
        gap = (in.x == PA_GAP_CODE || in.x == PA_OTHER_CODE);
        if (!miss && !gap) leaf_mul<R>(PT, in.x, a0, a1, a2, a3);
    } else {
        gap = (in.x == GAPT);
        if (!miss && !gap) {
            const V4<R> v = in.v;
            const V4<R> m0 = ld4<R>((const LDS V4<R> *)(PT + 0)), m1 = ld4<R>((const LDS V4<R> *)(PT + 4));
            const V4<R> m2 = ld4<R>((const LDS V4<R> *)(PT + 8)), m3 = ld4<R>((const LDS V4<R> *)(PT + 12));
            const R o0 = m0.x * v.x + m1.x * v.y + m2.x * v.z + m3.x * v.w;
            const R o1 = m0.y * v.x + m1.y * v.y + m2.y * v.z + m3.y * v.w;
            const R o2 = m0.z * v.x + m1.z * v.y + m2.z * v.z + m3.z * v.w;
            const R o3 = m0.w * v.x + m1.w * v.y + m2.w * v.z + m3.w * v.w;
            a0 *= o0; a1 *= o1; a2 *= o2; a3 *= o3;
            e += in.x;
        }
    }
}

}  // namespace pa
