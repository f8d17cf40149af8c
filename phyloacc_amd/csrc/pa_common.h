// pa_common.h -- plain structs shared by the host encoder and the device
// kernels: model constants, per-element descriptors and the node-table layout.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define PA_TAB __host__ __device__ inline
#else
#define PA_TAB inline
#endif

namespace pa {

constexpr int MAXN = 512;          // max nodes per tree (S <= 256)
constexpr int GAPT = -0x40000000;  // scale-exponent sentinel: subtree has only gap/N leaves

// node flags (per element, per node id)
enum : uint32_t {
    NF_MISSING = 1u,       // BPP_C::missing
    NF_IN_TREE = 2u,       // in BPP_C::nodes (the element's, possibly collapsed, tree)
    NF_UPPER = 4u,         // in BPP::upper (outgroup lineage)
    NF_UPPER_CONS = 8u,    // in BPP::upper_conserve
    NF_LEAF = 16u,
};

// Model constants passed by value to every kernel.
struct ModelDev {
    int S, N, I;                 // species, nodes, internal-node slots (N - S)
    int T, burnin, mcmc, adapt_freq, ropt;
    double cub, nlb;
    double cprior_a, cprior_b, nprior_a, nprior_b;
    double init_crate, init_nrate, init_grate, init_lrate, init_lrate2;
    double ha_g, hb_g, ha_l, hb_l, ha_l2, hb_l2;
    double constomis, nconstomis;
    double pi[4];
    double lam[4];               // eigenvalues of Q
    double V[16], Vi[16];        // P(t)[i][j] = sum_k V[i][k] e^{lam_k t} Vi[k][j] = P_{i->j}
    uint32_t key0, key1;         // SEED
    int r1_shared;
};

// Per-element descriptor.  Offsets index the engine's arenas.
struct ElemDev {
    int GG, P, Ppad, root;
    int n_nodes, n_int, n_up, n_down;   // n_up / n_down: number of levels
    int has_amb, pad0, pad1, pad2;
    int64_t off_codes;   // uint8 [S][Ppad] leaf codes, pattern index fastest
    int64_t off_w;       // int32 [P] pattern weights
    int64_t off_site;    // int32 [GG] site -> pattern
    int64_t off_tab;     // int32 node table (TAB_WORDS(N) words)
};

// Node table layout (int32 words, per element; N = model N).
//   ch0[N] ch1[N] par[N] flags[N] nodes[N] internal[N]
//   up[N] up_off[N+1] down[N] down_off[N+1] tz[N] dist2 (double, 2N words)
//   up: internal nodes of the element's tree by height (1, 2, ...);
//       level h occupies up[up_off[h-1] .. up_off[h]).
//   down: non-root nodes by depth (1, 2, ...), same convention.
//   tz[s]: position of node s in the reverse post-order (root = 0), the Z-update draw slot.
PA_TAB int tab_ch0(int N) { return 0; }
PA_TAB int tab_ch1(int N) { return N; }
PA_TAB int tab_par(int N) { return 2 * N; }
PA_TAB int tab_flags(int N) { return 3 * N; }
PA_TAB int tab_nodes(int N) { return 4 * N; }
PA_TAB int tab_int(int N) { return 5 * N; }
PA_TAB int tab_up(int N) { return 6 * N; }
PA_TAB int tab_upoff(int N) { return 7 * N; }
PA_TAB int tab_down(int N) { return 8 * N + 1; }
PA_TAB int tab_downoff(int N) { return 9 * N + 1; }
PA_TAB int tab_tz(int N) { return 10 * N + 2; }
PA_TAB int tab_dist(int N) { return (11 * N + 3) & ~1; }       // even -> 8-byte aligned doubles
PA_TAB int tab_words(int N) { return tab_dist(N) + 2 * N + 2; }  // even, so every table stays aligned

}  // namespace pa
