// pa_kernels.h -- launch interface between the engine host code and the
// HIP kernels in pa_engine.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "pa_common.h"

#define PA_GAP_CODE 4
#define PA_OTHER_CODE 5
#define PA_STATUS_NUMERIC_FAIL 2

namespace pa {

struct ChainArgs {
    const ElemDev *elems;
    const int32_t *work;      // element ids in schedule order
    int n_work;
    int *queue;               // work counter, reset before each launch
    const uint8_t *codes;     // arenas
    const int32_t *wts;
    const int32_t *site;
    const int32_t *tab;
    char *scratch;            // per-workgroup slots
    int64_t slot_bytes;
    int64_t off_parts, off_exps, off_plin, off_logt, off_amb;
    // outputs, indexed by work position
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
size_t slot_scratch_bytes(int N, int S, int Pcap, int fp64, size_t *off /*[5]*/);
hipError_t launch_kernel(const ModelDev &M, const ChainArgs &A, int fp64, int BS, int grid, int kind, hipStream_t st);
hipError_t launch_reset(int *q, hipStream_t st);

}  // namespace pa
