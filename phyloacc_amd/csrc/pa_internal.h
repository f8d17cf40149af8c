// pa_internal.h -- what the engine's other host modules may ask of an engine
// (pa_host.cpp owns the struct).
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "../../include/phyloacc_gpu.h"

namespace pa {

struct EngineView {
    int S, N;                // species and nodes of the model tree
    uint32_t key0, key1;     // SEED, the key of every random stream
};
void engine_view(const pa_engine *e, EngineView *v);
// records msg as the engine's last error and returns st
pa_status engine_fail(pa_engine *e, pa_status st, const std::string &msg);
// Element id must be an unfiltered element added with its own tree.  Writes
// the missing flag of every leaf (more than gap_prop of its sites are gaps)
// to leaf_missing [S], and fails unless the stored tree's internal nodes
// have the children in children [N*2] (either order) when that is given.
pa_status element_gene_tree(pa_engine *e, int32_t id, uint8_t *leaf_missing, const int32_t *children);

// Extra work of a gene-tree likelihood launch; arrays are per caller row
// (host memory unless marked device).
struct GtExtra {
    const double *d_len = nullptr;                    // device: the lengths [n][N] (the host lengths are then unused)
    double *d_row0 = nullptr, *d_row1 = nullptr;      // device: the log-likelihoods by row as well
    const int32_t *tabs = nullptr, *hdrs = nullptr;   // per-row trees [n][tab_words(N)], [n][4] (gene_table)
    int8_t *d_hist = nullptr;                         // device: the histories (row r at hist_off[r],
    const int64_t *hist_off = nullptr;                //   node v at v * hist_ns[r] from there)
    const int32_t *hist_ns = nullptr;
    const int32_t *hist_len = nullptr;                // draw mode: sites to draw per row
    int hist_draw = 0, hist_sweep = 0;
    const int32_t *cand_branch = nullptr, *cand_n = nullptr, *cand_node = nullptr, *cand_kind = nullptr,
                  *cand_sites = nullptr;
    const double *cand_len = nullptr;
    int cand_max = 0;
    double *cand_score = nullptr;                     // out: [n][cand_max]
    // resident partials: row r's at d_big + row_big_off[r] (loglik_resident_bytes),
    // node slots chosen by d_alt [n][N]; rows with d_path[r] >= 0 prune that
    // node's path to the root only (device arrays, rows in the caller's order)
    char *d_big = nullptr;
    const int64_t *row_big_off = nullptr;             // host
    const uint8_t *d_alt = nullptr;
    const int32_t *d_path = nullptr;
};
// The engine's gene-node histories: the device arena of the last
// pa_gt_sample_history call and which rows it holds.
struct GtHistory {
    int8_t *d = nullptr;
    size_t cap = 0;
    bool valid = false;
    std::vector<int32_t> ids, len;
    std::vector<int64_t> off;
};
GtHistory &engine_history(pa_engine *e);
// A resident-partials likelihood launch prepared once for a fixed batch
// (x: d_len, d_big, row_big_off, d_alt; d_row0/1 optional); each run uploads
// the prefix lengths (lens0 null: the pattern sum, no prefixes) and launches
// on the engine stream without waiting; gt_plan_stats waits and records the
// launch's time and counted bytes.
struct GtPlan;
pa_status gt_plan_build(pa_engine *e, const int32_t *ids, int32_t n, const double *piA, int32_t precision,
                        const GtExtra &x, GtPlan **out);
// same_lens: lens0/lens1 are those of the previous run, already on the
// device; add_count: the counted bytes add to the previous run's.
pa_status gt_plan_run(pa_engine *e, GtPlan *p, const int32_t *lens0, const int32_t *lens1, const int32_t *d_path,
                      bool same_lens = false, bool add_count = false);
pa_status gt_plan_stats(pa_engine *e, GtPlan *p);
void gt_plan_free(GtPlan *p);
// element id's padded pattern count (-1 for an unknown id)
int element_ppad(pa_engine *e, int32_t id);
// grows the history arena to at least `bytes` (device memory)
pa_status engine_history_reserve(pa_engine *e, size_t bytes);
// One gene-tree likelihood launch (pa_loglik_lengths_split's, lens0 may be
// null) with the extra work of x.
pa_status gt_launch(pa_engine *e, const int32_t *ids, int32_t n, const double *lengths, const double *piA,
                    const int32_t *lens0, const int32_t *lens1, int32_t precision, double *out0, double *out1,
                    const GtExtra &x);
// Node table of element id's leaves under the tree children [N*2], and its
// header: root, node count, height levels, depth levels.
pa_status gene_table(pa_engine *e, int32_t id, const int32_t *children, int32_t *T, int32_t *hdr);
// Makes the node table tabs[k] (with its header hdrs[k], both from
// gene_table) the stored tree of element ids[k] (host and, when uploaded,
// device).  The ids must be distinct: the elements are updated in parallel.
pa_status element_set_trees(pa_engine *e, const int32_t *ids, int n, const int32_t *const *tabs,
                            const int32_t *const *hdrs);
// The engine's stream (after the device is set up), for launches that a
// likelihood launch on it must follow.
pa_status engine_stream(pa_engine *e, hipStream_t *st);

}  // namespace pa
