// pa_gtdev.hip -- device-resident PhyloAcc-GT node-height moves
// (GTree::Sample_BranchLen, src/PhyloAcc-GT/genetree.cpp:1658-1965) for a
// batch of gene-tree elements whose trees stay in HBM between moves.
//
// A proposal kernel (one wave per element) draws each element's new height
// and computes the proposed tree's branch lengths, coalescent prior and
// missing-gene term with the host path's own gene-tree functions
// (pa_gtcore.h); the likelihood launch reads those lengths in place; a
// decision kernel (one lane per element) accepts or rejects and updates the
// chain state.  No host threads and no PCIe round trip per move.
//
// No multiply-add contraction in this file: the proposal heights, lengths
// and priors round every product and sum the way the host path does, so the
// two paths keep the same trees bit for bit.
#pragma clang fp contract(off)
#include "pa_gtdev.h"
#include "pa_rng.h"

namespace pa {

// One 64-lane workgroup (one wave) per element; the element's rows and the
// species tree are copied into LDS.  Lane 0 takes the move's few scalar
// steps (the window, the draws, the moved node's new species); the rest runs
// across the lanes, with every sum still taken in the host's order so both
// paths give the same bits:
// - the prior (GTree::priorTree): gene nodes born per species by LDS atomics;
//   the lineages entering species x, the leaves below x less the gene nodes
//   born in internal species strictly below it; each born node's rank in its
//   species' height order, and the height before it, by ranking; each
//   coalescence term and each species' tail written at its place in the
//   host's sequence (species post-order, born nodes by height, then the tail;
//   +0.0 where the host adds nothing, which leaves a sum unchanged);
// - the missing-gene term: a term per gene node at its gene post-order place;
// - lane 0 adds the prior sequence while lane 1 adds the missing-gene one;
// - the branch lengths, one node per lane.
__global__ __launch_bounds__(64) void gt_hprop_kernel(GtPropArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int status_sh, total_sh;
    __shared__ double lc[4];
    const int i = blockIdx.x, t = threadIdx.x;
    const int N = a.N, S = a.S;
    double *H = (double *)lds, *theta = H + N, *rate = theta + N, *ldens = rate + N, *h = ldens + N;
    double *seq = h + N, *lzseq = seq + 2 * N;
    int32_t *up = (int32_t *)(lzseq + N), *post = up + N, *ipost = post + N, *nleaf = ipost + N, *ch = nleaf + N;
    int32_t *par = ch + 2 * N, *Z = par + N, *where = Z + N, *end = where + N, *cnt = end + N, *below = cnt + N;
    int32_t *soff = below + N, *fill = soff + N + 1, *blist = fill + N;
    uint8_t *miss = (uint8_t *)(blist + 2 * N);
    const size_t r = (size_t)i * N;
    for (int k = t; k < N; k += 64) {
        H[k] = a.sp.H[k];
        theta[k] = a.sp.theta[k];
        rate[k] = a.rate[k];
        ldens[k] = a.ldens[k];
        up[k] = a.sp.up[k];
        const int pk = a.sp.post[k];
        post[k] = pk;
        ipost[pk] = k;
        nleaf[k] = a.nleaf[k];
        h[k] = a.h[r + k];
        par[k] = a.par[r + k];
        Z[k] = a.Z[r + k];
        where[k] = a.where[r + k];
        miss[k] = a.miss[r + k];
        cnt[k] = 0;
        below[k] = 0;
        fill[k] = 0;
    }
    for (int k = t; k < 2 * N; k += 64) ch[k] = a.ch[2 * r + k];
    if (t < 4) lc[t] = a.lc[t];
    const int root = a.root[i], v = a.gnode[i];
    __syncthreads();
    for (int k = t; k < N; k += 64) end[k] = k == root ? N - 1 : where[par[k]];
    __syncthreads();
    const gt::SpView sp = {S, N, up, H, theta, post, nullptr, nullptr};
    gt::TreeView g = {S, N, root, ch, par, nullptr, h, where, end};
    if (t == 0 && a.lengths_only) {
        status_sh = 1;
    } else if (t == 0) {
        const gt::Window w = gt::window(sp, g, v, a.delta);
        a.win[i] = w;
        status_sh = w.status;
        if (w.status != 0) {
            // no proposal: the launch scores the current tree, and the decision skips it
            a.prop[i] = h[v];
        } else {
            const Key key{a.key0, a.key1};
            const uint32_t c3 = tagword(0, TAG_GT_BRANCH), c2 = (uint32_t)a.m;
            const double u = slot_uniform(key, (uint32_t)a.slot, c2, c3, 0);
            const double prop = fmax(u, 1e-5) * (w.upper - w.lower) + w.lower;
            a.prop[i] = prop;
            a.rdraw[i] = slot_uniform(key, (uint32_t)a.slot, c2, c3, 1);
            h[v] = prop;
            gt::repoint(sp, g, v);
            a.wnew[i] = where[v];
        }
    }
    __syncthreads();
    if (status_sh == 0) {
        for (int u = S + t; u < N; u += 64) {
            const int x = where[u];
            atomicAdd(&cnt[x], 1);
            if (x >= S)
                for (int z = up[x]; z < N; z = up[z]) atomicAdd(&below[z], 1);
        }
        __syncthreads();
        {
            // soff: each species' start in the prior sequence (species
            // post-order; cnt + 1 places per internal species), chunks per
            // lane and a shuffle scan
            const int c = (N + 63) / 64, lo = min(N, t * c), hi = min(N, lo + c);
            int sum = 0;
            for (int q = lo; q < hi; q++) sum += post[q] >= S ? cnt[post[q]] + 1 : 0;
            int inc = sum;
            for (int d = 1; d < 64; d <<= 1) {
                const int y = __shfl_up(inc, d, 64);
                if (t >= d) inc += y;
            }
            int run = inc - sum;
            for (int q = lo; q < hi; q++) {
                soff[q] = run;
                run += post[q] >= S ? cnt[post[q]] + 1 : 0;
            }
            if (t == 63) total_sh = inc;
        }
        __syncthreads();
        // each internal species' born nodes as a list (in any order) at its
        // sequence place
        for (int u = S + t; u < N; u += 64) {
            const int x = where[u];
            if (x >= S) blist[soff[ipost[x]] + atomicAdd(&fill[x], 1)] = u;
        }
        __syncthreads();
        // born nodes: rank in the species' height order (ties by index), the
        // height before, the term; the last one also writes the tail
        for (int u = S + t; u < N; u += 64) {
            const int x = where[u];
            if (x < S) continue;
            const double hu = h[u];
            const int base = soff[ipost[x]], ke = nleaf[x] - below[x], n_b = cnt[x];
            int rk = 0;
            double prev = H[x];
            for (int j = 0; j < n_b; j++) {
                const int q = blist[base + j];
                const double hq = h[q];
                if (hq < hu || (hq == hu && q < u)) {
                    rk++;
                    prev = fmax(prev, hq);
                }
            }
            seq[base + rk] = ke == 0 ? 0.0 : gt::coal_term(ldens[x], rate[x], hu, prev, ke - rk);
            if (rk == n_b - 1) {
                const int kf = ke - n_b;
                seq[base + n_b] =
                    ke != 0 && kf > 1 && up[x] != N ? -gt::coal_tail(rate[x], H[up[x]], hu, kf) : 0.0;
            }
        }
        // internal species with no born node: the tail alone
        for (int x = S + t; x < N; x += 64) {
            if (cnt[x] != 0) continue;
            const int ke = nleaf[x] - below[x];
            seq[soff[ipost[x]]] = ke != 0 && ke > 1 && up[x] != N ? -gt::coal_tail(rate[x], H[up[x]], H[x], ke) : 0.0;
        }
        for (int q = t; q < N; q += 64) {
            double tv = 0;
            lzseq[q] = gt::logpz_term(sp, g, miss, Z, lc, a.order[r + q], &tv) ? tv : 0.0;
        }
        __syncthreads();
        if (t < 2) {
            const double *arr = t == 0 ? seq : lzseq;
            const int len = t == 0 ? total_sh : N;
            double acc = 0;
#pragma unroll 8
            for (int j = 0; j < len; j++) acc += arr[j];
            if (t == 0) a.prior_new[i] = acc;
            else a.lz_new[i] = acc;
        }
    }
    const double r1 = a.rates[2 * i], r2 = a.rates[2 * i + 1];
    for (int k = t; k < N; k += 64) a.len[r + k] = gt::branch_length(sp, g, k, Z, r1, r2);
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
    a.where[(size_t)i * a.N + a.gnode[i]] = a.wnew[i];
    if (a.alt) {
        const int32_t *par = a.par + (size_t)i * a.N;
        uint8_t *alt = a.alt + (size_t)i * a.N;
        for (int x = a.gnode[i], k = 0; x >= 0 && x < a.N && k < a.N; x = par[x], k++) {
            alt[x] ^= 1;
            if (x == a.root[i]) break;
        }
    }
    a.loglik[2 * i] = lnew[0];
    a.loglik[2 * i + 1] = lnew[1];
    if (a.logpz[2 * i + 1] == 0) a.logpz[2 * i] = a.lz_new[i];
    else a.logpz[2 * i] = a.logpz[2 * i + 1] = a.lz_new[i];
    a.prior[i] = a.prior_new[i];
}

__global__ __launch_bounds__(256) void gt_scatter_tables_kernel(const int32_t *src, int W, const int64_t *dst_off,
                                                                 int32_t *d_tab, const ElemDev *rec,
                                                                 const int32_t *dst_id, ElemDev *d_elems)
{
    const int k = blockIdx.x;
    int32_t *dst = d_tab + dst_off[k];
    const int32_t *s = src + (size_t)k * W;
    for (int w = threadIdx.x; w < W; w += blockDim.x) dst[w] = s[w];
    constexpr int RW = (int)(sizeof(ElemDev) / 4);
    if (threadIdx.x < RW)
        ((int32_t *)(d_elems + dst_id[k]))[threadIdx.x] = ((const int32_t *)(rec + k))[threadIdx.x];
}

hipError_t launch_gt_scatter_tables(const int32_t *src, int W, const int64_t *dst_off, int32_t *d_tab,
                                    const ElemDev *rec, const int32_t *dst_id, ElemDev *d_elems, int n,
                                    hipStream_t st)
{
    static_assert(sizeof(ElemDev) % 4 == 0 && sizeof(ElemDev) / 4 <= 256, "element record as words");
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gt_scatter_tables_kernel, dim3(n), dim3(256), 0, st, src, W, dst_off, d_tab, rec, dst_id,
                       d_elems);
    return hipGetLastError();
}

hipError_t launch_gt_hprop(const GtPropArgs &a, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
    const size_t lds = gt_hprop_lds(a.N);
    if (lds > GT_HPROP_LDS_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gt_hprop_kernel, dim3(a.n), dim3(64), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_gt_hdecide(const GtDecideArgs &a, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gt_hdecide_kernel, dim3((a.n + 127) / 128), dim3(128), 0, st, a);
    return hipGetLastError();
}

}  // namespace pa
