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
// chain failure and warning bits, as PA_FLAG_* in include/phyloacc_gpu.h
#define PA_F_NAN 1
#define PA_F_POS_INF 2
#define PA_F_NEG_INF 4
#define PA_F_SAMPLER 8
#define PA_F_PROPOSAL_RANGE 16

namespace pa {

// Arguments of the log-likelihood launch.  Work position w evaluates element
// work[w] either at ancestral states Zin[w] and rates rates_in[w] over the
// element's stored branch lengths, or, when len_in is set, at the per-call
// lengths (row len_row[w], or w, of len_in) with every branch at rate 1 (Zin,
// rates_in unused).
struct LoglikArgs {
    const ElemDev *elems;
    const int32_t *work;      // element ids in schedule order
    int n_work;
    int *queue;               // work counter, reset before each launch
    const uint8_t *codes;     // arenas
    const int32_t *wts;
    const int32_t *tab;
    int n_elems;              // elements in the arenas: ids outside [0, n_elems) score NaN
    char *scratch;            // one slot per workgroup
    int64_t slot_bytes;
    // Elements too long for a slot: work position w keeps its partials at
    // big + big_off[w] (loglik_slot_bytes of its own Ppad); big_off[w] < 0,
    // or big_off null, means the workgroup's slot.
    char *big;
    const int64_t *big_off;
    // Resident partials (gene-tree height moves), or null: every work
    // position's partials stay at big + big_off[w] (loglik_resident_bytes of
    // its Ppad: two slots per internal node), node x of row r in its second
    // slot when alt[r][x] != 0.  A row with path_node[r] >= 0 prunes only the
    // path from that node to the root, into the slots alt does not select,
    // and reads every other node's partial where it is.
    const uint8_t *alt;
    const int32_t *path_node;
    const int32_t *Zin;       // [n_work][N] (when len_in is null)
    const double *rates_in;   // [n_work][2] (r1, r2) (when len_in is null)
    const double *len_in;     // [rows][N] per-call branch lengths, or null
    const int32_t *len_row;   // [n_work] row of len_in for each work position, or null (row w)
    double *ll_out;           // [n_work]
    // gBGC substitution model, or null: branch s of work position w takes
    // P(dist) of class k = c(Z) + 3 * (Bin[w][s] != 0), c(0) = 0 (background),
    // c(2) = 2 (accelerated), else 1 (conserved), whose eigensystem is
    // gb_eig[w][k] = {lam[4], V[16], Vi[16]} (gbgc_eig_kernel).
    const double *gb_eig;     // [n_work][6][GB_EIG_WORDS]
    const int32_t *Bin;       // [n_work][N]
    unsigned long long *counters;   // [0] += algorithmic HBM bytes of the launch, or null
    // Per-element substitution model for the own-tree path, or null: work
    // position w prunes with the eigensystem el_eig[w] = {lam[4], V[16],
    // Vi[16]} and sums its root with the base frequencies el_pi[w][4]
    // (gt_eig_kernel).
    const double *el_eig;     // [n_work][GB_EIG_WORDS]
    const double *el_pi;      // [n_work][4]
    // Two site prefixes per caller row, or null: instead of the pattern sum,
    // the root sums run over the element's sites, g < pre_len[r][0] into
    // ll_out and g < pre_len[r][1] into ll_out2 (r = len_row[w], or w; the
    // site -> pattern map is `site`).
    const int32_t *pre_len;   // [rows][2]
    double *ll_out2;          // [n_work]
    // The log-likelihoods by caller row as well, or null (device-resident
    // callers): row_out0[r] = ll_out, row_out1[r] = ll_out2.
    double *row_out0, *row_out1;
    // PhyloAcc-GT gene-tree moves.  Arrays below are per caller row: work
    // position w uses row len_row[w] (or w).
    // Per-call trees, or null: row r prunes the tree of node table
    // tab_call[r] (tab_words(N) words; children, parents, flags, post-order,
    // levels) with root, node count and level count in hdr_call[r][0..2].
    const int32_t *tab_call;
    const int32_t *hdr_call;  // [rows][4]
    const int32_t *site;      // the site -> pattern arena (ElemDev::off_site)
    // Gene-node histories, or null: row r's lives at hist + hist_off[r],
    // node v's sites at v * hist_ns[r] from there (int8: leaves 0-3, -1
    // ambiguous, 4 no base; internal nodes their drawn base, 4 when missing).
    // hist_draw: after pruning, draw sites g < hist_len[r] there (tag
    // GT_HIST, sub-stream g, sweep hist_sweep, slot v); else the history is
    // only read, by the candidate scores.
    int8_t *hist;
    const int64_t *hist_off;
    const int32_t *hist_ns, *hist_len;
    int hist_draw, hist_sweep;
    // Candidate scores, or null: for row r, branch cand_branch[r] and
    // candidates q < cand_n[r] (node cand_node, kind cand_kind, distance
    // cand_len; cand_max per row), sum over sites g < cand_sites[r] the score
    // of kind 0 (node's base, or the pi-average when it has none), 1 (the
    // pi-average) or 2 (the branch's own base's pi-row) into cand_score,
    // reading the history.
    const int32_t *cand_branch, *cand_n, *cand_node, *cand_kind, *cand_sites;
    const double *cand_len;
    int cand_max;
    double *cand_score;
};

constexpr int GB_CLASSES = 6, GB_EIG_WORDS = 36;

// Eigensystems of the six gBGC classes of every work position: class k has
// selection 0, sel[w][0] (conserved) or sel[w][1] (accelerated) for k % 3 =
// 0, 1, 2, and bias sel[w][2] when k >= 3, else 0 (Q from Q0 by gbgc_rate).
// A class whose Q is not time-reversible gets NaN eigenvalues.
hipError_t launch_gbgc_eig(const double *Q0 /*[16], host*/, const double *sel /*[n][3], device*/, int n,
                           double *eig /*[n][6][36], device*/, hipStream_t st);

// Eigensystem and base frequencies of PhyloAcc-GT's strand-symmetric model
// for every work position: pi = (a, 1/2 - a, 1/2 - a, a) with a = piA[w],
// Q from the six exchangeabilities inst (gt_strand_q).  A non-reversible
// result gets NaN eigenvalues.
hipError_t launch_gt_eig(const double *inst /*[6], host*/, const double *piA /*[n], device*/, int n,
                         double *eig /*[n][36], device*/, double *pi /*[n][4], device*/, hipStream_t st);

// LDS bytes of one workgroup (dynamic arrays + context + argument blocks).
size_t loglik_smem_bytes(int N, int fp64);
// Bytes of one workgroup's HBM slot for elements of up to Pcap (padded)
// patterns: the partials [I][Pcap] (4 values each) and their exponents.
size_t loglik_slot_bytes(int N, int S, int Pcap, int fp64);
// Bytes of one element's resident partials at Ppad patterns (two slots per
// internal node and their exponents).
size_t loglik_resident_bytes(int N, int S, int Ppad, int fp64);
// Dynamic LDS bytes of one workgroup (the launch's dynamic allocation).
size_t loglik_dyn_smem_bytes(int N, int fp64);
// Workgroups of the likelihood kernel resident on one CU (registers, LDS, launch bound).
int loglik_blocks_per_cu(int fp64, int BS, int N);
hipError_t launch_loglik(const ModelDev &M, const LoglikArgs &A, int fp64, int BS, int grid, hipStream_t st);

// Size of the device scratch array of the allocation-free likelihood entry
// point, and the device address of the current device's copy (looked up once
// per device with that device current).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)256 << 20;
hipError_t noalloc_scratch(int device, char **ptr);
hipError_t launch_reset(int *q, hipStream_t st);

}  // namespace pa
