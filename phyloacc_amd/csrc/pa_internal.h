// pa_internal.h -- what the engine's other host modules may ask of an engine
// (pa_host.cpp owns the struct).
#pragma once
#include <stdint.h>
#include <string>
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

}  // namespace pa
