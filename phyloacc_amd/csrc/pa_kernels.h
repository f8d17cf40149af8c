// pa_kernels.h -- launch interface between the engine host code and the
// log-likelihood kernel in pa_engine.hip (the chain kernel: pa_chain.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "pa_common.h"

#define PA_GAP_CODE 4
#define PA_OTHER_CODE 5
#define PA_STATUS_NUMERIC_FAIL 2

namespace pa {

// Chain state between launches of a resumable run.  The partials, exponents
// and class tables stay in the record's scratch slot; everything else that a
// sweep carries to the next one is here.
struct ChainState {
    int32_t ms;               // model slot being run (0..2), 3 when all three are done
    int32_t m;                // next iteration of that model; -1: not initialised yet
    int32_t accept_n, accept_c, max_m, status, pad0, pad1;
    double loglik, nrate, crate, g, l, l2, prop_n, prop_c, max_loglik;
    unsigned long long bytes;
    int8_t Z[MAXN], maxZ[MAXN];
    uint8_t cur[MAXN], fixZ[MAXN];
};

struct ChainArgs {
    const ElemDev *elems;
    const int32_t *work;      // element ids in schedule order
    int n_work;
    int *queue;               // work counter, reset before each launch
    const uint8_t *codes;     // arenas
    const int32_t *wts;
    const int32_t *site;
    const int32_t *tab;
    char *scratch;            // scratch slots
    int64_t slot_bytes;       // one slot per workgroup (when slot_off is null)
    const int64_t *slot_off;  // [records] byte offset of each record's own slot, or null
    const int32_t *rec;       // [n_work] record of each work position, or null (record = position)
    ChainState *state;        // [records] resumable chains, or null (run to the end)
    int32_t *done;            // [records] set to 1 when all three models finished (resumable runs)
    int sweep_budget;         // sweeps per work item in this launch (resumable runs)
    const int *n_work_dev;    // when set, the number of work items is read here at launch
    // outputs, indexed by record
    double *res_d;            // [n_work][3][8]
    int32_t *res_i;           // [n_work][3][4]
    int8_t *maxZ;             // [n_work][3][N]
    double *tr_d;             // [n_work][3][7][T] or null
    int8_t *tr_Z;             // [n_work][3][T][N] or null
    // fixed-(Z, r) log-likelihood
    const int32_t *Zin;       // [n_work][N]
    const double *rates_in;   // [n_work][2] (r1, r2)
    double *ll_out;           // [n_work]
    const double *len_in;     // [n_work][N] per-call branch lengths (own-tree elements) or null
    unsigned long long *counters;   // [0]: algorithmic HBM bytes of the launch (or null)
};

size_t chain_smem_bytes(int N, int S, int BS, int fp64);
// Byte layout of one scratch slot for an element with Ppad patterns:
// off = {parts, exps, plin, logt, amb}; returns the slot size.
__host__ __device__ inline size_t slot_layout(int N, int S, int Ppad, int fp64, size_t *off)
{
    const int I = N - S;
    const size_t rs = fp64 ? 8 : 4;
    size_t o = 0;
    off[0] = o; o = (o + (size_t)2 * I * Ppad * 4 * rs + 255) & ~(size_t)255;   // parts
    off[1] = o; o = (o + (size_t)2 * I * Ppad * 4 + 255) & ~(size_t)255;        // exps
    off[2] = o; o = (o + (size_t)3 * N * 16 * 8 + 255) & ~(size_t)255;          // plin
    off[3] = o; o = (o + (size_t)3 * N * 16 * 8 + 255) & ~(size_t)255;          // logt
    off[4] = o; o = (o + (size_t)S * 24 * 4 + 255) & ~(size_t)255;              // amb
    return o;
}
hipError_t launch_loglik(const ModelDev &M, const ChainArgs &A, int fp64, int BS, int grid, hipStream_t st);
hipError_t launch_reset(int *q, hipStream_t st);

}  // namespace pa
