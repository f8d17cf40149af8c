// pa_gtdev.hip -- device-resident PhyloAcc-GT node-height moves
// (GTree::Sample_BranchLen, src/PhyloAcc-GT/genetree.cpp:1658-1965) for a
// batch of gene-tree elements whose trees stay in HBM between moves.
//
// One lane per element runs the move's scalar logic -- the embedding of its
// gene tree in the species tree, the window, the draws, the proposal's branch
// lengths, coalescent prior and missing-gene term -- with the same code the
// host path runs (pa_gtcore.h), on the element's own workspace rows.  The
// work is pointer-chasing over a few hundred nodes, not a dense contraction:
// its place is beside the likelihood launch, not in it, and it takes the host
// threads and the PCIe round trip of every move out of the sweep.
//
// No multiply-add contraction in this file: the proposal heights, lengths
// and priors round every product and sum the way the host path does, so the
// two paths keep the same trees bit for bit.
#pragma clang fp contract(off)
#include "pa_gtdev.h"
#include "pa_rng.h"

namespace pa {

__global__ void gt_hprop_kernel(GtPropArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int N = a.N, S = a.S;
    int32_t *where = a.ws_i + (size_t)i * (5 * N + 1);
    int32_t *end = where + N, *cnt = end + N, *off = cnt + N, *leave = off + N + 1;
    double *hs = a.ws_d + (size_t)i * N;
    uint8_t *miss = a.ws_u + (size_t)i * N;
    double *h = a.h + (size_t)i * N;
    const int32_t *Z = a.Z + (size_t)i * N;
    const double r1 = a.rates[2 * i], r2 = a.rates[2 * i + 1];
    gt::TreeView g = {S, N, a.root[i], a.ch + (size_t)i * 2 * N, a.par + (size_t)i * N, a.order + (size_t)i * N, h,
                      where, end};
    gt::embed(a.sp, g);
    const int v = a.gnode[i];
    const gt::Window w = gt::window(a.sp, g, v, a.delta);
    a.win[i] = w;
    double *len = a.len + (size_t)i * N;
    if (w.status != 0) {
        // no proposal: the launch scores the current tree, and the decision skips it
        gt::lengths(a.sp, g, Z, r1, r2, len);
        a.prop[i] = h[v];
        return;
    }
    const Key key{a.key0, a.key1};
    const uint32_t c3 = tagword(0, TAG_GT_BRANCH), c2 = (uint32_t)a.m;
    const double u = slot_uniform(key, (uint32_t)a.slot, c2, c3, 0);
    const double prop = fmax(u, 1e-5) * (w.upper - w.lower) + w.lower;
    a.prop[i] = prop;
    a.rdraw[i] = slot_uniform(key, (uint32_t)a.slot, c2, c3, 1);
    const double old = h[v];
    h[v] = prop;
    gt::repoint(a.sp, g, v);
    gt::lengths(a.sp, g, Z, r1, r2, len);
    a.prior_new[i] = gt::prior(a.sp, g, cnt, off, leave, hs);
    gt::missing(g, a.lm + (size_t)i * S, miss);
    a.lz_new[i] = gt::logpz(a.sp, g, miss, Z, a.c2m, a.n2m);
    h[v] = old;
}

__global__ void gt_hdecide_kernel(GtDecideArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    a.accepted[i] = 0;
    const gt::Window w = a.win[i];
    if (w.status != 0) return;
    const int ind = a.indicator[i];
    const double lnew[2] = {a.lnew0[i], a.lnew1[i]};
    const double d = gt::height_ratio(w, a.prop[i], ind, lnew, a.loglik + 2 * i, a.logpz + 2 * i, a.lz_new[i],
                                      a.prior_new[i], a.prior[i]);
    if (!(log(a.rdraw[i]) < d)) return;
    a.accepted[i] = 1;
    a.h[(size_t)i * a.N + a.gnode[i]] = a.prop[i];
    a.loglik[2 * i] = lnew[0];
    a.loglik[2 * i + 1] = lnew[1];
    if (a.logpz[2 * i + 1] == 0) a.logpz[2 * i] = a.lz_new[i];
    else a.logpz[2 * i] = a.logpz[2 * i + 1] = a.lz_new[i];
    a.prior[i] = a.prior_new[i];
}

hipError_t launch_gt_hprop(const GtPropArgs &a, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gt_hprop_kernel, dim3((a.n + 63) / 64), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_gt_hdecide(const GtDecideArgs &a, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gt_hdecide_kernel, dim3((a.n + 127) / 128), dim3(128), 0, st, a);
    return hipGetLastError();
}

}  // namespace pa
