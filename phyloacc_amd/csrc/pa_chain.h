// pa_chain.h -- launch interface of the chain kernel in pa_chain.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "pa_common.h"
#include "pa_kernels.h"

namespace pa {
namespace chain {

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
    unsigned long long *counters;   // [0]: algorithmic HBM bytes of the launch (or null)
};

// LDS bytes of one workgroup (dynamic arrays + context + model constants).
size_t chain_smem_bytes(int N, int S, int BS, int fp64);
// Byte layout of one workgroup's HBM slot for elements of up to Pcap
// (padded) patterns: off = {parts, exps, plin, logt, amb}; returns its size.
size_t slot_scratch_bytes(int N, int S, int Pcap, int fp64, size_t *off /*[5]*/);
hipError_t launch_chain(const ModelDev &M, const ChainArgs &A, int fp64, int BS, int grid, hipStream_t st);
hipError_t launch_reset(int *q, hipStream_t st);

}  // namespace chain
}  // namespace pa
