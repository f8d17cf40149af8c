// pa_gtdev.h -- launch interface of the device-resident gene-tree height
// moves (pa_gtdev.hip): one wave per element proposes, a path likelihood
// launch over the resident partials scores the proposals, one lane per
// element decides.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pa_gtcore.h"
#include "pa_common.h"

namespace pa {

// Per element i (rows of N nodes): the gene tree with its embedding (the
// species of every gene node, kept current by the decisions), its missing
// flags, species-branch classes and rates, the node to move, and outputs.
// Per species: the leaves below it, 1 / theta and log(2 / theta); lc: the
// missing-gene logarithms (log(1 - c2m), log(c2m), log(1 - n2m), log(n2m)).
// The proposal works on a copy of the element's rows in LDS.
struct GtPropArgs {
    gt::SpView sp;               // device arrays (up, H, theta, post)
    const int32_t *nleaf;
    const double *rate, *ldens;
    double lc[4];
    int n, S, N;
    const int32_t *ch, *par, *order, *root, *where;
    const double *h;             // [n][N]
    const uint8_t *miss;         // [n][N]
    const int32_t *Z;
    const double *rates;         // [n][2] (r1 conserved, r2 accelerated)
    const int32_t *gnode;        // [n]
    double delta;
    uint32_t key0, key1;
    int m, slot;
    int lengths_only;            // 1: no move, the current trees' lengths only
    double *len;                 // [n][N] out: the proposal's branch lengths (the current ones for an empty window)
    gt::Window *win;             // [n] out
    double *prop, *rdraw, *prior_new, *lz_new;   // [n] out
    int32_t *wnew;               // [n] out: the moved node's species under the proposal
};

struct GtDecideArgs {
    int n, N;
    const gt::Window *win;
    const double *prop, *rdraw, *prior_new, *lz_new;
    const int32_t *wnew;
    const double *lnew0, *lnew1;  // [n] the proposals' prefix log-likelihoods
    const int32_t *gnode, *indicator;
    double *h, *loglik, *logpz, *prior;   // the chain state, updated on acceptance
    int32_t *where;
    // resident partials: on acceptance the nodes from gnode to the root
    // switch slots (the launch wrote their new partials into the other one)
    const int32_t *par, *root;
    uint8_t *alt;                // [n][N]
    int32_t *accepted;
};

// LDS bytes of one proposal workgroup at N species-tree nodes (the species
// tree, the element's rows and the move's workspace)
inline size_t gt_hprop_lds(int N) { return (size_t)129 * N + 8; }
constexpr size_t GT_HPROP_LDS_MAX = 64 * 1024 - 64;   // and 48 bytes of static LDS

hipError_t launch_gt_hprop(const GtPropArgs &a, hipStream_t st);

// Element k's new node table (W words at src + k W) to d_tab + dst_off[k],
// its record to d_elems[dst_id[k]]: one workgroup per element.
hipError_t launch_gt_scatter_tables(const int32_t *src, int W, const int64_t *dst_off, int32_t *d_tab,
                                    const ElemDev *rec, const int32_t *dst_id, ElemDev *d_elems, int n,
                                    hipStream_t st);
hipError_t launch_gt_hdecide(const GtDecideArgs &a, hipStream_t st);

}  // namespace pa
