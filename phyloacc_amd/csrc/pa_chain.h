// pa_chain.h -- launch interface of the chain kernel in pa_chain.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "pa_common.h"
#include "pa_kernels.h"

namespace pa {
namespace chain {

// Per (work position, model) result words: RES_D doubles (max_loglik, final
// nrate, crate, g, l, l2) and RES_I ints (max_m, status, accept_n,
// accept_c, flags).
constexpr int RES_D = 8, RES_I = 8;

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
    int64_t off_plin, off_logt, off_amb, off_parts;   // within a slot (slot_scratch_bytes)
    // Elements too long for a slot's partial region: work position w keeps its
    // partials at big + big_off[w] (chain_part_bytes of its own Ppad);
    // big_off[w] < 0, or big_off null, means the workgroup's slot.
    char *big;
    const int64_t *big_off;
    // outputs, indexed by work position
    double *res_d;            // [n_work][3][RES_D]
    int32_t *res_i;           // [n_work][3][RES_I]
    int8_t *maxZ;             // [n_work][3][N]
    double *tr_d;             // [n_work][3][7][T] or null
    int8_t *tr_Z;             // [n_work][3][T][N] or null
    unsigned long long *counters;   // [0]: algorithmic HBM bytes of the launch (or null)
    // Per-element completion flags [n_work] in host-visible memory, or null:
    // done[w] becomes 1 once work position w's three result rows are written
    // (res_d and res_i then live in host-visible memory too).
    int32_t *done;
};

// LDS bytes of one workgroup (dynamic arrays + context + model constants).
size_t chain_smem_bytes(int N, int S, int BS, int fp64);
// Dynamic LDS bytes of one workgroup (the launch's dynamic allocation).
size_t chain_dyn_smem_bytes(int N, int S, int BS, int fp64);
// Workgroups of the chain kernel resident on one CU (registers, LDS, launch bound).
int chain_blocks_per_cu(int fp64, int BS, int N, int S, int shape = 0);
// Pattern-loop shapes of the chain kernel (nodes per pruning round, nodes
// per sampling chunk), 0 = the default; see Shape in pa_chain.hip.
constexpr int N_SHAPES = 3;
// Partials and exponents of one element of Ppad padded patterns: two buffers
// per internal node; the exponents start at chain_exps_off.
size_t chain_part_bytes(int N, int S, int Ppad, int fp64);
// Byte layout of one workgroup's HBM slot whose partial region holds elements
// of up to Pcap (padded) patterns: off = {plin, logt, amb, parts}; returns its size.
size_t slot_scratch_bytes(int N, int S, int Pcap, int fp64, size_t *off /*[4]*/);
hipError_t launch_chain(const ModelDev &M, const ChainArgs &A, int fp64, int BS, int grid, hipStream_t st,
                        int shape = 0);
hipError_t launch_reset(int *q, hipStream_t st);
// Zeroes `words` 32-bit words at p with a kernel (graph-capturable).
hipError_t launch_clear(void *p, int words, hipStream_t st);

}  // namespace chain
}  // namespace pa
