// pa_gtdev.h -- launch interface of the device-resident gene-tree height
// moves (pa_gtdev.hip): one lane per element proposes, a likelihood launch
// scores the proposals, one lane per element decides.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pa_gtcore.h"

namespace pa {

// Per element i (rows of N nodes): the gene tree, its species-branch
// classes and rates, leaf missing flags, the node to move; per-element
// workspace (5N + 1 ints, N doubles, N bytes) and outputs.
struct GtPropArgs {
    gt::SpView sp;               // device arrays
    int n, S, N;
    const int32_t *ch, *par, *order, *root;
    double *h;                   // [n][N]: written for a moment, restored
    const int32_t *Z;
    const double *rates;         // [n][2] (r1 conserved, r2 accelerated)
    const uint8_t *lm;           // [n][S]
    const int32_t *gnode;        // [n]
    double delta, c2m, n2m;
    uint32_t key0, key1;
    int m, slot;
    int32_t *ws_i;
    double *ws_d;
    uint8_t *ws_u;
    double *len;                 // [n][N] out: the proposal's branch lengths (the current ones for an empty window)
    gt::Window *win;             // [n] out
    double *prop, *rdraw, *prior_new, *lz_new;   // [n] out
};

struct GtDecideArgs {
    int n, N;
    const gt::Window *win;
    const double *prop, *rdraw, *prior_new, *lz_new;
    const double *lnew0, *lnew1;  // [n] the proposals' prefix log-likelihoods
    const int32_t *gnode, *indicator;
    double *h, *loglik, *logpz, *prior;   // the chain state, updated on acceptance
    int32_t *accepted;
};

hipError_t launch_gt_hprop(const GtPropArgs &a, hipStream_t st);
hipError_t launch_gt_hdecide(const GtDecideArgs &a, hipStream_t st);

}  // namespace pa
