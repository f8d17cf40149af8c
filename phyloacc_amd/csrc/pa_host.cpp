// pa_host.cpp -- host side of the engine: .mod parsing, Q eigensystem,
// element encoding and pattern compression, device arenas, and the C ABI
// declared in include/phyloacc_gpu.h.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cctype>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/phyloacc_gpu.h"
#include "pa_common.h"
#include "pa_kernels.h"
#include "pa_chain.h"
#include "pa_rng.h"
#include "pa_eig.h"
#include "pa_internal.h"
#include "pa_gtdev.h"

using namespace pa;

static_assert(PA_F_NAN == PA_FLAG_NAN && PA_F_POS_INF == PA_FLAG_POS_INF && PA_F_NEG_INF == PA_FLAG_NEG_INF &&
                  PA_F_SAMPLER == PA_FLAG_SAMPLER && PA_F_PROPOSAL_RANGE == PA_FLAG_PROPOSAL_RANGE,
              "device failure bits match the C ABI");

// ------------------------------------------------------------- .mod parse

namespace {

struct PNode {
    std::string name;
    double dist = 0;
    int kid[2] = {-1, -1};
    int nk = 0;
    int id = -1;
};

// Recursive-descent Newick reader with the reference's conventions
// (src/PhyloAcc-ST/newick.cpp:145-331): a leaf's length follows its last ':';
// an internal label sits between ')' and ':'; lengths are read as float.
struct Newick {
    const std::string &s;
    std::vector<PNode> nodes;
    bool bad = false;
    explicit Newick(const std::string &t) : s(t) {}

    int parse(size_t a, size_t b)
    {
        int me = (int)nodes.size();
        nodes.emplace_back();
        if (a >= b) { bad = true; return me; }
        if (s[a] != '(') {
            size_t colon = std::string::npos;
            for (size_t i = a; i < b; i++) if (s[i] == ':') colon = i;
            size_t e = colon == std::string::npos ? b : colon;
            nodes[me].name = s.substr(a, e - a);
            nodes[me].dist = colon == std::string::npos ? 0.0 : (double)(float)atof(s.c_str() + colon + 1);
            return me;
        }
        int depth = 0;
        size_t start = a + 1, i = a;
        std::vector<int> kids;
        for (; i < b; i++) {
            char ch = s[i];
            if (ch == '(') depth++;
            else if (ch == ')') { if (--depth == 0) break; }
            else if (ch == ',' && depth == 1) { kids.push_back(parse(start, i)); start = i + 1; }
        }
        if (i >= b) { bad = true; return me; }
        kids.push_back(parse(start, i));
        if (kids.size() != 2) { bad = true; return me; }
        nodes[me].kid[0] = kids[0];
        nodes[me].kid[1] = kids[1];
        nodes[me].nk = 2;
        size_t j = i + 1, colon = std::string::npos, e = j;
        while (e < b && s[e] != ';' && s[e] != ':') e++;
        if (e < b && s[e] == ':') colon = e;
        nodes[me].name = s.substr(j, e - j);
        if (colon != std::string::npos) nodes[me].dist = (double)(float)atof(s.c_str() + colon + 1);
        return me;
    }
};

std::string trim3(const std::string &x)
{
    size_t a = x.find_first_not_of(" \t\n");
    if (a == std::string::npos) return "";
    size_t b = x.find_last_not_of(" \t\n");
    return x.substr(a, b - a + 1);
}

std::vector<std::string> tokens(const std::string &x)
{
    std::vector<std::string> out;
    size_t i = 0;
    while (i < x.size()) {
        while (i < x.size() && (x[i] == ' ' || x[i] == '\t')) i++;
        size_t j = i;
        while (j < x.size() && x[j] != ' ' && x[j] != '\t') j++;
        if (j > i) out.push_back(x.substr(i, j - i));
        i = j;
    }
    return out;
}

}  // namespace

extern "C" pa_status pa_parse_mod(const char *text, int32_t max_nodes, int32_t *S_out, int32_t *children,
                                  int32_t *parent, double *dist, double *pi, double *Q, char *names)
{
    if (!text || !S_out) return PA_E_INPUT;
    std::vector<std::string> lines;
    {
        const char *p = text;
        while (true) {
            const char *e = strchr(p, '\n');
            if (!e) { lines.emplace_back(p); break; }
            lines.emplace_back(p, e - p);
            p = e + 1;
        }
    }
    std::string tree;
    double pib[4] = {0, 0, 0, 0}, Qb[16] = {0};
    for (size_t li = 0; li < lines.size(); li++) {
        if (lines[li].empty()) break;   // LoadPhyloTree stops at the first empty line
        std::vector<std::string> tk = tokens(trim3(lines[li]));
        if (tk.empty()) continue;
        if (tk[0] == "BACKGROUND:") {
            for (size_t g = 1; g < tk.size() && g <= 4; g++) pib[g - 1] = atof(tk[g].c_str());
        } else if (tk[0] == "RATE_MAT:") {
            for (int i = 0; i < 4 && li + 1 < lines.size(); i++) {
                std::vector<std::string> row = tokens(trim3(lines[++li]));
                for (size_t g = 0; g < row.size() && g < 4; g++) Qb[i * 4 + g] = atof(row[g].c_str());
            }
        } else if (tk[0] == "TREE:" && tk.size() > 1) {
            tree = tk[1];
        }
    }
    if (tree.empty()) return PA_E_INPUT;
    size_t end = tree.size();
    while (end > 0 && tree[end - 1] == ';') end--;
    Newick nw(tree);
    int root = nw.parse(0, end);
    if (nw.bad) return PA_E_INPUT;
    for (auto &n : nw.nodes)
        if (!std::isfinite(n.dist)) return PA_E_INPUT;   // lengths below 1e-10 (negative too) act as 0
    int S = 0;
    for (auto &n : nw.nodes) S += n.nk == 0;
    int N = 2 * S - 1;
    if ((int)nw.nodes.size() != N || N > max_nodes || N > MAXN) return PA_E_INPUT;
    // TravelTree2 numbering: post-order, leaves and internal nodes counted apart
    int lid = 0, iid = 0;
    std::vector<int> stack{root}, post;
    std::vector<int> state(nw.nodes.size(), 0);
    while (!stack.empty()) {
        int v = stack.back();
        PNode &n = nw.nodes[v];
        if (state[v] < n.nk) { stack.push_back(n.kid[state[v]++]); continue; }
        stack.pop_back();
        n.id = n.nk == 0 ? lid++ : S + iid++;
    }
    *S_out = S;
    for (int i = 0; i < N; i++) {
        if (children) { children[2 * i] = children[2 * i + 1] = -1; }
        if (parent) parent[i] = -1;
    }
    for (auto &n : nw.nodes) {
        if (dist) dist[n.id] = n.dist;
        if (names) {
            memset(names + 64 * n.id, 0, 64);
            strncpy(names + 64 * n.id, n.name.c_str(), 63);
        }
        if (n.nk == 2) {
            int a = nw.nodes[n.kid[0]].id, b = nw.nodes[n.kid[1]].id;
            if (a > b) std::swap(a, b);
            if (children) { children[2 * n.id] = a; children[2 * n.id + 1] = b; }
            if (parent) { parent[a] = n.id; parent[b] = n.id; }
        }
    }
    if (pi) memcpy(pi, pib, sizeof pib);
    if (Q) memcpy(Q, Qb, sizeof Qb);
    return PA_OK;
}

// ------------------------------------------------------------- eigensystem

namespace {

// Q = V diag(lam) Vi for a time-reversible Q (BPP ctor eig_gen + inv,
// src/PhyloAcc-ST/bpp.hpp:357-362), via the symmetric form D^1/2 Q D^-1/2.
bool eigensystem(const double *Q, double *lam, double *V, double *Vi, std::string &err)
{
    switch (reversible_eig4(Q, lam, V, Vi)) {
    case EIG_OK: return true;
    case EIG_NO_STATIONARY: err = "RATE_MAT has no positive stationary distribution"; return false;
    case EIG_NOT_REVERSIBLE: err = "RATE_MAT is not time-reversible"; return false;
    default: err = "eigendecomposition of RATE_MAT failed"; return false;
    }
}

}  // namespace

// ---------------------------------------------------------------- encoder

namespace {

struct HostElem {
    int filtered = 0, GG = 0, P = 0, Ppad = 0, root = -1, has_amb = 0;
    int n_up = 0, n_down = 0;
    std::vector<int32_t> tab;         // node table (tab_words(N))
    std::vector<uint8_t> codes;       // [S][Ppad]
    std::vector<int32_t> w, site;     // [P], [GG]
    std::vector<uint8_t> site_codes;  // [GG][S]
    std::vector<int> nodes, internal;
    // gene trees: the stored children of each internal node (smaller id
    // first) and the leaves' missing flags, kept compact for the check that
    // a caller's tree is the stored one
    std::vector<int16_t> kids;        // [2 (N - S)]
    std::vector<uint8_t> leaf_miss;   // [S]
};

// The compact form of a gene tree's internal nodes' children (smaller id first).
static void compact_kids(const int32_t *ch0, const int32_t *ch1, int S, int N, std::vector<int16_t> &kids)
{
    kids.resize(2 * (size_t)(N - S));
    for (int v = S; v < N; v++) {
        const int a = ch0[v], b = ch1[v];
        kids[2 * (v - S)] = (int16_t)std::min(a, b);
        kids[2 * (v - S) + 1] = (int16_t)std::max(a, b);
    }
}

inline uint8_t code_of(int y, int gapchar)
{
    // BPP_C constructor switch (src/PhyloAcc-ST/bpp_c.hpp:220-261)
    switch (y) {
    case 'a': return PA_CODE_A;
    case 'c': return PA_CODE_C;
    case 'g': return PA_CODE_G;
    case 't': return PA_CODE_T;
    case 'r': return PA_CODE_R;
    case 'y': return PA_CODE_Y;
    case 'k': return PA_CODE_K;
    case 'm': return PA_CODE_M;
    case 's': return PA_CODE_S;
    case 'w': return PA_CODE_W;
    default: return y == gapchar ? PA_CODE_GAP : PA_CODE_OTHER;
    }
}

struct TreeModel {
    int S, N;
    std::vector<int> ch, par;
    std::vector<double> dist;
    std::vector<uint8_t> upper, upper_cons;
    std::vector<int> conserve;
};

struct Collapse {
    std::vector<int> &ch2, &par2;
    std::vector<double> &dist2;
    const std::vector<uint8_t> &missing, &upper;
    int S, &root;
    // BPP_C::getSubtree_missing (src/PhyloAcc-ST/bpp_c.cpp:64-113)
    void run(int ss, int child)
    {
        if (!upper[ss] || ss < S) return;
        int s = ss;
        while (upper[s] && s >= S) {
            int a = ch2[2 * s], b = ch2[2 * s + 1];
            int keep;
            if (missing[a] && !missing[b]) { par2[a] = -1; keep = b; }
            else if (!missing[a] && missing[b]) { par2[b] = -1; keep = a; }
            else { run(a, 0); run(b, 1); return; }
            dist2[keep] += dist2[s];
            if (s == root) root = keep;
            else { ch2[2 * par2[ss] + child] = keep; par2[keep] = par2[ss]; }
            s = keep;
        }
    }
};

// Node table of one element's tree (the layout of pa_common.h): children
// and parents (the root's parent N), flags, the post-order of the nodes
// under root (BPP_C::getSubtree, src/PhyloAcc-ST/bpp_c.cpp:44-60), the
// internal nodes by height level, the non-root nodes by depth level, the
// Z-draw slots and the branch lengths.  own: a gene tree (no outgroup flags).
// The table into T (tab_words(N) words, every word written); the post-order
// and internal-node lists stay in per-thread storage behind nodes_out /
// internal_out until the thread's next call.
static void fill_table(const TreeModel &tm, const std::vector<int> &ch2, const std::vector<int> &par2,
                       const double *dist2, const std::vector<uint8_t> &missing, bool own, int root, int32_t *T,
                       int &n_up, int &n_down, const std::vector<int> **nodes_out, const std::vector<int> **internal_out)
{
    const int S = tm.S, N = tm.N;
    // scratch kept per thread: tables are rebuilt for every proposed gene tree
    thread_local std::vector<int> nodes, internal, height, depth, cnt;
    thread_local std::vector<std::pair<int, int>> st;
    // post-order of the element's tree (BPP_C::getSubtree, bpp_c.cpp:44-60)
    nodes.clear();
    {
        st.assign(1, {root, 0});
        while (!st.empty()) {
            auto &[v, k] = st.back();
            if (v >= S && k < 2) { int x = ch2[2 * v + k]; k++; if (x != -1) st.push_back({x, 0}); continue; }
            nodes.push_back(v);
            st.pop_back();
        }
    }
    internal.clear();
    for (int v : nodes) if (v >= S) internal.push_back(v);
    // levels for the Z forward-backward
    height.assign(N, 0);
    depth.assign(N, 0);
    int hmax = 0, dmax = 0;
    for (int v : nodes)
        if (v >= S) { height[v] = 1 + std::max(height[ch2[2 * v]], height[ch2[2 * v + 1]]); hmax = std::max(hmax, height[v]); }
    for (int k = (int)nodes.size() - 1; k >= 0; k--) {
        int v = nodes[k];
        depth[v] = v == root ? 0 : depth[par2[v]] + 1;
        dmax = std::max(dmax, depth[v]);
    }
    memset(T, 0, sizeof(int32_t) * tab_words(N));
    for (int i = 0; i < N; i++) {
        T[tab_ch0(N) + i] = ch2[2 * i];
        T[tab_ch1(N) + i] = ch2[2 * i + 1];
        T[tab_par(N) + i] = par2[i];
        T[tab_flags(N) + i] = (missing[i] ? NF_MISSING : 0) | (!own && tm.upper[i] ? NF_UPPER : 0) |
                              (!own && tm.upper_cons[i] ? NF_UPPER_CONS : 0) | (i < S ? NF_LEAF : 0);
    }
    for (int v : nodes) T[tab_flags(N) + v] |= NF_IN_TREE;
    for (size_t k = 0; k < nodes.size(); k++) {
        T[tab_nodes(N) + k] = nodes[k];
        T[tab_tz(N) + nodes[k]] = (int)nodes.size() - 1 - (int)k;
    }
    for (size_t k = 0; k < internal.size(); k++) T[tab_int(N) + k] = internal[k];
    // up: internal nodes by height (post-order within a level); down: non-root
    // nodes by depth (reverse post-order within a level); stable counting sorts
    {
        cnt.assign(std::max(hmax, dmax) + 2, 0);
        for (int v : internal) cnt[height[v]]++;
        int q = 0;
        for (int h = 1; h <= hmax; h++) {
            const int c = cnt[h];
            cnt[h] = q;
            q += c;
            T[tab_upoff(N) + h - 1] = q;
        }
        for (int v : internal) T[tab_up(N) + cnt[height[v]]++] = v;
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int v : nodes) if (depth[v] >= 1) cnt[depth[v]]++;
        q = 0;
        for (int d = 1; d <= dmax; d++) {
            const int c = cnt[d];
            cnt[d] = q;
            q += c;
            T[tab_downoff(N) + d - 1] = q;
        }
        for (int k = (int)nodes.size() - 1; k >= 0; k--)
            if (depth[nodes[k]] >= 1) T[tab_down(N) + cnt[depth[nodes[k]]]++] = nodes[k];
    }
    double *D = (double *)(T + tab_dist(N));
    for (int i = 0; i < N; i++) D[i] = dist2[i];
    n_up = hmax;
    n_down = dmax;
    *nodes_out = &nodes;
    *internal_out = &internal;
}

void fill_tree_table(const TreeModel &tm, const std::vector<int> &ch2, const std::vector<int> &par2,
                     const std::vector<double> &dist2, const std::vector<uint8_t> &missing, bool own, int root,
                     HostElem &E)
{
    E.tab.resize(tab_words(tm.N));
    const std::vector<int> *nodes, *internal;
    fill_table(tm, ch2, par2, dist2.data(), missing, own, root, E.tab.data(), E.n_up, E.n_down, &nodes, &internal);
    E.root = root;
    E.nodes = *nodes;
    E.internal = *internal;
}

void encode_element(const TreeModel &tm, const pa_options &o, const char *const *rows, int64_t start, int len,
                    HostElem &E, const int32_t *own_children = nullptr)
{
    const int S = tm.S, N = tm.N;
    E = HostElem();
    if (len < o.min_len) { E.filtered = 1; return; }
    // site codes [len][S]
    std::vector<uint8_t> sc((size_t)len * S);
    for (int s = 0; s < S; s++) {
        const char *r = rows[s] + start;
        for (int g = 0; g < len; g++) sc[(size_t)g * S + s] = code_of(tolower((unsigned char)r[g]), o.gapchar);
    }
    // column trimming (bpp_c.hpp:268-301)
    std::vector<int> keep;
    keep.reserve(len);
    for (int g = 0; g < len; g++) {
        if (o.trim_gap_percent < 1) {
            int mis = 0;
            for (int s = 0; s < S; s++) mis += sc[(size_t)g * S + s] == PA_CODE_GAP || sc[(size_t)g * S + s] == PA_CODE_OTHER;
            if (mis > S * o.trim_gap_percent) continue;
        }
        keep.push_back(g);
    }
    if (o.trim_gap_percent < 1 && (int)keep.size() < o.min_len) { E.filtered = 1; return; }
    const int GG = (int)keep.size();
    E.GG = GG;
    E.site_codes.resize((size_t)GG * S);
    for (int g = 0; g < GG; g++) memcpy(&E.site_codes[(size_t)g * S], &sc[(size_t)keep[g] * S], S);
    // gap counts and missing species (bpp_c.hpp:304-351)
    std::vector<int> ngap(S, 0);
    for (int g = 0; g < GG; g++)
        for (int s = 0; s < S; s++) ngap[s] += E.site_codes[(size_t)g * S + s] == PA_CODE_GAP;
    std::vector<uint8_t> missing(N, 0);
    std::vector<int> ch2(tm.ch), par2(tm.par);
    std::vector<double> dist2(tm.dist);
    if (own_children) {
        // an element with its own tree (a gene tree over the same leaves): lengths come per call
        ch2.assign(own_children, own_children + 2 * N);
        for (int i = 0; i < N; i++) par2[i] = -1;
        for (int i = S; i < N; i++) { par2[ch2[2 * i]] = i; par2[ch2[2 * i + 1]] = i; }
        std::fill(dist2.begin(), dist2.end(), 0.0);
    }
    for (int s = S; s < N; s++) {
        for (int q = 0; q < 2; q++) {
            int x = ch2[2 * s + q];
            if (x < S && ngap[x] > o.gap_prop * GG) missing[x] = 1;
        }
        if (missing[ch2[2 * s]] && missing[ch2[2 * s + 1]]) missing[s] = 1;
    }
    // CONSERVE filter (bpp_c.hpp:361-373)
    int ct = 0;
    for (int x : tm.conserve) ct += missing[x];
    if (!own_children && ct > o.conserve_prop * (double)tm.conserve.size()) { E.filtered = 1; return; }
    // outgroup-lineage (or whole-tree) collapse (bpp_c.hpp:379-390)
    int root = N - 1;
    std::vector<uint8_t> all(N, 1), none(N, 0);   // gene trees are never collapsed
    Collapse col{ch2, par2, dist2, missing, own_children ? none : (o.prune_tree ? all : tm.upper), S, root};
    col.run(N - 1, -1);
    par2[root] = N;
    if (root < S) { E.filtered = 1; return; }   // nothing left to prune over
    fill_tree_table(tm, ch2, par2, dist2, missing, own_children != nullptr, root, E);
    if (own_children) {
        const int32_t *T = E.tab.data();
        compact_kids(T + tab_ch0(N), T + tab_ch1(N), S, N, E.kids);
        E.leaf_miss.resize(S);
        for (int q = 0; q < S; q++) E.leaf_miss[q] = (T[tab_flags(N) + q] & NF_MISSING) ? 1 : 0;
    }
    // pattern compression (first-appearance order) or one pattern per site
    E.site.resize(GG);
    std::vector<int> first;
    if (o.compress_patterns) {
        std::unordered_map<std::string, int> map;
        map.reserve(GG * 2);
        for (int g = 0; g < GG; g++) {
            std::string key((const char *)&E.site_codes[(size_t)g * S], S);
            auto it = map.find(key);
            if (it == map.end()) {
                int id = (int)first.size();
                map.emplace(std::move(key), id);
                first.push_back(g);
                E.site[g] = id;
                E.w.push_back(1);
            } else {
                E.site[g] = it->second;
                E.w[it->second]++;
            }
        }
    } else {
        for (int g = 0; g < GG; g++) { E.site[g] = g; first.push_back(g); E.w.push_back(1); }
    }
    E.P = (int)first.size();
    E.Ppad = (E.P + 63) & ~63;
    E.codes.assign((size_t)S * E.Ppad, PA_CODE_GAP);
    for (int p = 0; p < E.P; p++)
        for (int s = 0; s < S; s++) {
            uint8_t cd = E.site_codes[(size_t)first[p] * S + s];
            E.codes[(size_t)s * E.Ppad + p] = cd;
            if (cd >= PA_CODE_R) E.has_amb = 1;
        }
}

}  // namespace

// ----------------------------------------------------------------- engine

struct pa_engine {
    int device = 0;
    pa_options opt;
    ModelDev M;
    double Q0[16];                  // RATE_MAT (the gBGC classes are built from it)
    TreeModel tm;
    std::string err;
    std::mutex err_mu;
    // the last batch's schedule and likelihood scratch plan: a batch that is
    // launched again (an MCMC loop over the same elements) reuses them
    std::mutex sched_mu;
    std::vector<int32_t> sched_ids, sched_order;
    std::vector<HostElem> elems;
    bool dev_dirty = true;
    std::vector<struct pa_batch *> batches;   // prepared batches, freed with the engine
    hipEvent_t ll_t0 = nullptr, ll_t1 = nullptr;   // around the last log-likelihood launch
    pa_run_stats ll_stats{};
    // device arenas
    ElemDev *d_elems = nullptr;
    uint8_t *d_codes = nullptr;
    int32_t *d_w = nullptr, *d_site = nullptr, *d_tab = nullptr;
    int *d_queue = nullptr;
    hipStream_t stream = nullptr;
    int n_cu = 0;
    size_t lds_per_cu = 0;
    pa_tuning tune = {0, 0, 0, 0, 0.0};     // batched chain launches: block size (0: by mean pattern count), shape
    uint64_t arena_gen = 0;                  // bumped whenever the device arenas move (prepared batches check it)
    // log-likelihood calls: scratch slots and per-call branch lengths, kept
    // between calls and grown when a call needs more (freed with the engine)
    void *d_ll_scratch = nullptr, *d_ll_len = nullptr;
    size_t ll_scratch_bytes = 0, ll_len_bytes = 0;
    void *h_ll_len = nullptr;       // pinned staging of the per-call lengths
    size_t h_ll_len_bytes = 0;
    void *h_tab_stage = nullptr;    // pinned staging of accepted gene trees' node tables
    size_t h_tab_stage_bytes = 0;
    size_t ll_last_plan_bytes = 0;           // scratch the last likelihood launch used
    // the allocation-free likelihood entry point: this device's copy of its
    // scratch array, and the arenas' longest element and mean pattern count
    char *d_noalloc = nullptr;
    int max_ppad = 0;
    double mean_p = 0;
    std::vector<ElemDev> dev_elems;          // host copy of the uploaded descriptors (tree updates patch both)
    GtHistory hist;                          // gene-node histories of the last history draw (device)
};

static pa_status fail(pa_engine *e, pa_status st, const std::string &msg)
{
    if (e) {
        std::lock_guard<std::mutex> lock(e->err_mu);   // host phases may run on several threads
        e->err = msg;
    }
    return st;
}

#define HIPCHK(e, x)                                                                      \
    do {                                                                                  \
        hipError_t _r = (x);                                                              \
        if (_r != hipSuccess)                                                             \
            return fail(e, _r == hipErrorOutOfMemory ? PA_E_OOM : PA_E_DEVICE,            \
                        std::string(#x) + ": " + hipGetErrorString(_r));                  \
    } while (0)

extern "C" void pa_default_options(pa_options *o)
{
    memset(o, 0, sizeof *o);
    o->burnin = 200; o->mcmc = 800; o->adapt_freq = 500;
    o->init_crate = 0.5; o->init_nrate = 1.0;
    o->cprior_a = 5; o->cprior_b = 0.04; o->nprior_a = 10; o->nprior_b = 0.2;
    o->ropt = 1; o->cub = 1; o->nlb = 0.6;
    o->init_grate = 0.5; o->init_lrate = 0.3; o->init_lrate2 = 0.1;
    o->hyper_grate_a = 3; o->hyper_grate_b = 1; o->hyper_lrate_a = 1; o->hyper_lrate_b = 9;
    o->hyper_lrate2_a = 1; o->hyper_lrate2_b = 9;
    o->gap_prop = 0.8; o->conserve_prop = 0.8; o->constomis = 0.01; o->trim_gap_percent = 1;
    o->min_len = 50; o->prune_tree = 0; o->gapchar = '-'; o->seed = 5;
    o->r1_shared_uniform = 1; o->compress_patterns = 1;
}

extern "C" const char *pa_last_error(const pa_engine *e) { return e ? e->err.c_str() : "null engine"; }

extern "C" int32_t pa_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" double pa_rng_uniform(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, int32_t lane)
{
    U4 v = philox(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
    return u01(pick(v, lane & 3));
}

extern "C" pa_status pa_engine_create(const pa_model *model, const pa_options *opts, int32_t device, pa_engine **out)
{
    if (!model || !out) return PA_E_INPUT;
    *out = nullptr;
    const int S = model->S, N = 2 * S - 1;
    if (S < 2 || N > MAXN) return PA_E_INPUT;
    pa_engine *e = new pa_engine();
    if (opts) e->opt = *opts; else pa_default_options(&e->opt);
    const pa_options &o = e->opt;
    if (o.burnin < 0 || o.mcmc < 1 || o.adapt_freq < 1) { delete e; return PA_E_INPUT; }
    TreeModel &tm = e->tm;
    tm.S = S; tm.N = N;
    tm.ch.assign(model->children, model->children + 2 * N);
    tm.par.assign(model->parent, model->parent + N);
    tm.dist.assign(model->dist, model->dist + N);
    tm.upper.assign(N, 0);
    tm.upper_cons.assign(N, 0);
    // the tree must be the reference numbering: binary, post-order ids, root N-1
    for (int i = 0; i < N; i++) {
        bool leaf = i < S;
        if (leaf != (tm.ch[2 * i] == -1)) { e->err = "children[] do not match leaves 0..S-1"; *out = e; return PA_E_INPUT; }
        if (!leaf && (tm.ch[2 * i] >= i || tm.ch[2 * i + 1] >= i)) { e->err = "internal ids are not post-order"; *out = e; return PA_E_INPUT; }
    }
    // BPP::getUppertree (src/PhyloAcc-ST/bpp.cpp:486-504)
    auto upper = [&](const int32_t *sp, int n, std::vector<uint8_t> &set) -> bool {
        for (int i = 0; i < n; i++) {
            int p = sp[i];
            if (p < 0 || p >= S) return false;
            while (p != N - 1) { set[p] = 1; p = tm.par[p]; }
        }
        set[N - 1] = 1;
        return true;
    };
    if (!upper(model->outgroup, model->n_outgroup, tm.upper) || !upper(model->conserve, model->n_conserve, tm.upper_cons)) {
        e->err = "OUTGROUP/CONSERVE species id out of range";
        *out = e;
        return PA_E_INPUT;
    }
    tm.conserve.assign(model->conserve, model->conserve + model->n_conserve);
    ModelDev &M = e->M;
    memset(&M, 0, sizeof M);
    M.S = S; M.N = N; M.I = N - S;
    M.T = o.burnin + o.mcmc; M.burnin = o.burnin; M.mcmc = o.mcmc; M.adapt_freq = o.adapt_freq; M.ropt = o.ropt;
    M.cub = o.cub; M.nlb = o.nlb;
    M.cprior_a = o.cprior_a; M.cprior_b = o.cprior_b; M.nprior_a = o.nprior_a; M.nprior_b = o.nprior_b;
    M.init_crate = o.init_crate; M.init_nrate = o.init_nrate; M.init_grate = o.init_grate;
    M.init_lrate = o.init_lrate; M.init_lrate2 = o.init_lrate2;
    M.ha_g = o.hyper_grate_a; M.hb_g = o.hyper_grate_b; M.ha_l = o.hyper_lrate_a; M.hb_l = o.hyper_lrate_b;
    M.ha_l2 = o.hyper_lrate2_a; M.hb_l2 = o.hyper_lrate2_b;
    M.constomis = o.constomis; M.nconstomis = 1.0;
    for (int i = 0; i < 4; i++) M.pi[i] = model->pi[i];
    M.key0 = (uint32_t)o.seed; M.key1 = (uint32_t)(o.seed >> 32);
    M.r1_shared = o.r1_shared_uniform;
    std::string err;
    if (!eigensystem(model->Q, M.lam, M.V, M.Vi, err)) { e->err = err; *out = e; return PA_E_NUMERIC; }
    memcpy(e->Q0, model->Q, sizeof e->Q0);
    e->device = device;
    *out = e;
    return PA_OK;
}

// Device set-up happens on first use, so that encoding and inspection work
// on hosts without a GPU.
static pa_status ensure_device(pa_engine *e)
{
    if (e->stream) return PA_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= e->device) return fail(e, PA_E_DEVICE, "no HIP device");
    HIPCHK(e, hipSetDevice(e->device));
    hipDeviceProp_t prop;
    HIPCHK(e, hipGetDeviceProperties(&prop, e->device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(e, PA_E_DEVICE, std::string("engine is built for gfx950, device is ") + prop.gcnArchName);
    e->n_cu = prop.multiProcessorCount;
    e->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
    HIPCHK(e, hipStreamCreate(&e->stream));
    HIPCHK(e, hipMalloc(&e->d_queue, 256));
    HIPCHK(e, hipEventCreate(&e->ll_t0));
    HIPCHK(e, hipEventCreate(&e->ll_t1));
    HIPCHK(e, noalloc_scratch(e->device, &e->d_noalloc));
    return PA_OK;
}

static void free_arenas(pa_engine *e)
{
    (void)hipFree(e->d_elems); (void)hipFree(e->d_codes); (void)hipFree(e->d_w);
    (void)hipFree(e->d_site); (void)hipFree(e->d_tab);
    e->d_elems = nullptr; e->d_codes = nullptr; e->d_w = e->d_site = e->d_tab = nullptr;
}

static void batch_free(pa_batch *b);

extern "C" void pa_engine_destroy(pa_engine *e)
{
    if (!e) return;
    while (!e->batches.empty()) batch_free(e->batches.back());
    if (e->stream) {
        (void)hipSetDevice(e->device);
        (void)hipStreamSynchronize(e->stream);
        free_arenas(e);
        (void)hipFree(e->d_queue);
        (void)hipFree(e->d_ll_scratch);
        (void)hipFree(e->d_ll_len);
        (void)hipHostFree(e->h_ll_len);
        (void)hipHostFree(e->h_tab_stage);
        (void)hipFree(e->hist.d);
        if (e->ll_t0) (void)hipEventDestroy(e->ll_t0);
        if (e->ll_t1) (void)hipEventDestroy(e->ll_t1);
        (void)hipStreamDestroy(e->stream);
    }
    delete e;
}

extern "C" pa_status pa_elements_add(pa_engine *e, const char *const *rows, const int64_t *starts, const int32_t *lengths,
                                     int32_t n, int32_t *ids, int32_t *status)
{
    if (!e || !rows || !starts || !lengths || n < 0) return fail(e, PA_E_INPUT, "null argument");
    const size_t base = e->elems.size();
    e->elems.resize(base + n);
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 64) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
            for (int i = (int)t; i < n; i += (int)nt) encode_element(e->tm, e->opt, rows, starts[i], lengths[i], e->elems[base + i]);
        });
    for (auto &x : th) x.join();
    for (int i = 0; i < n; i++) {
        if (ids) ids[i] = (int32_t)(base + i);
        if (status) status[i] = e->elems[base + i].filtered ? PA_ELEM_FILTERED : PA_ELEM_OK;
    }
    e->dev_dirty = true;
    return PA_OK;
}

// Elements that carry their own tree over the same leaves (gene trees,
// PhyloAcc-GT): children is [n][N][2] in post-order numbering, root N-1.
extern "C" pa_status pa_elements_add_trees(pa_engine *e, const char *const *rows, const int64_t *starts,
                                           const int32_t *lengths, int32_t n, const int32_t *children, int32_t *ids,
                                           int32_t *status)
{
    if (!e || !rows || !starts || !lengths || !children || n < 0) return fail(e, PA_E_INPUT, "null argument");
    const int N = e->tm.N, S = e->tm.S;
    for (int i = 0; i < n; i++) {
        const int32_t *ch = children + (size_t)i * 2 * N;
        for (int v = 0; v < N; v++) {
            const bool leaf = v < S;
            if (leaf != (ch[2 * v] == -1) || (!leaf && (ch[2 * v] >= v || ch[2 * v + 1] >= v || ch[2 * v] < 0)))
                return fail(e, PA_E_INPUT, "element tree " + std::to_string(i) + " is not in post-order numbering");
        }
    }
    const size_t base = e->elems.size();
    e->elems.resize(base + n);
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 64) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
            for (int i = (int)t; i < n; i += (int)nt)
                encode_element(e->tm, e->opt, rows, starts[i], lengths[i], e->elems[base + i], children + (size_t)i * 2 * N);
        });
    for (auto &x : th) x.join();
    for (int i = 0; i < n; i++) {
        if (ids) ids[i] = (int32_t)(base + i);
        if (status) status[i] = e->elems[base + i].filtered ? PA_ELEM_FILTERED : PA_ELEM_OK;
    }
    e->dev_dirty = true;
    return PA_OK;
}

// Substitution length of every gene branch from the pieces it spends in
// species branches: sum of piece length x rate of the piece's species branch
// (Z = 0 / 1 / 2 -> 1 / r1 / r2); pieces shorter than 1e-8 add nothing, as the
// identity matrix of getlogTMc (src/PhyloAcc-GT/bpp.cpp:221-235).
extern "C" pa_status pa_gene_lengths(int32_t N, int32_t K, const int32_t *seg_gene, const int32_t *seg_species,
                                     const double *seg_len, const int32_t *Z, double r1, double r2, double *out)
{
    if (N <= 0 || K < 0 || !out) return PA_E_INPUT;
    for (int i = 0; i < N; i++) out[i] = 0;
    for (int k = 0; k < K; k++) {
        if (seg_gene[k] < 0 || seg_gene[k] >= N || seg_species[k] < 0) return PA_E_INPUT;
        if (seg_len[k] < 1e-8) continue;
        const int z = Z[seg_species[k]];
        out[seg_gene[k]] += seg_len[k] * (z == 2 ? r2 : (z == 1 ? r1 : 1.0));
    }
    return PA_OK;
}

extern "C" pa_status pa_element_get(const pa_engine *e, int32_t id, pa_element_info *info, int32_t *children2,
                                    int32_t *parent2, double *dist2, int32_t *missing, int32_t *nodes,
                                    uint8_t *site_codes, int32_t *site_pattern, int32_t *weights)
{
    if (!e || id < 0 || id >= (int)e->elems.size()) return PA_E_INPUT;
    const HostElem &E = e->elems[id];
    const int N = e->tm.N, S = e->tm.S;
    if (info) {
        info->filtered = E.filtered; info->GG = E.GG; info->P = E.P; info->root = E.root;
        info->n_nodes = (int)E.nodes.size(); info->n_internal = (int)E.internal.size();
    }
    if (E.filtered) return PA_OK;
    const int32_t *T = E.tab.data();
    for (int i = 0; i < N; i++) {
        if (children2) { children2[2 * i] = T[tab_ch0(N) + i]; children2[2 * i + 1] = T[tab_ch1(N) + i]; }
        if (parent2) parent2[i] = T[tab_par(N) + i];
        if (dist2) dist2[i] = ((const double *)(T + tab_dist(N)))[i];
        if (missing) missing[i] = (T[tab_flags(N) + i] & NF_MISSING) ? 1 : 0;
    }
    if (nodes) for (size_t k = 0; k < E.nodes.size(); k++) nodes[k] = E.nodes[k];
    if (site_codes) memcpy(site_codes, E.site_codes.data(), E.site_codes.size());
    if (site_pattern) memcpy(site_pattern, E.site.data(), E.site.size() * 4);
    if (weights) memcpy(weights, E.w.data(), E.w.size() * 4);
    (void)S;
    return PA_OK;
}

namespace pa {

void engine_view(const pa_engine *e, EngineView *v)
{
    v->S = e->tm.S; v->N = e->tm.N;
    v->key0 = e->M.key0; v->key1 = e->M.key1;
}

pa_status engine_fail(pa_engine *e, pa_status st, const std::string &msg) { return fail(e, st, msg); }

pa_status element_gene_tree(pa_engine *e, int32_t id, uint8_t *leaf_missing, const int32_t *children)
{
    if (id < 0 || id >= (int)e->elems.size()) return fail(e, PA_E_INPUT, "element id out of range");
    const HostElem &E = e->elems[id];
    if (E.filtered) return fail(e, PA_E_INPUT, "element " + std::to_string(id) + " is filtered");
    const int N = e->tm.N, S = e->tm.S;
    const int32_t *T = E.tab.data();
    const bool compact = (int)E.kids.size() == 2 * (N - S) && (int)E.leaf_miss.size() == S;
    if (compact) memcpy(leaf_missing, E.leaf_miss.data(), S);
    else
        for (int s = 0; s < S; s++) leaf_missing[s] = (T[tab_flags(N) + s] & NF_MISSING) ? 1 : 0;
    if (children)
        for (int v = S; v < N; v++) {
            const int x = children[2 * v], y = children[2 * v + 1];
            int a, b;
            if (compact) {
                a = E.kids[2 * (v - S)];
                b = E.kids[2 * (v - S) + 1];
            } else {
                a = T[tab_ch0(N) + v];
                b = T[tab_ch1(N) + v];
            }
            if (!((a == x && b == y) || (a == y && b == x)))
                return fail(e, PA_E_INPUT, "element " + std::to_string(id) + ": gene node " + std::to_string(v) +
                                               " has other children than the element's stored tree");
        }
    return PA_OK;
}

}  // namespace pa

// Upload every element's encoding into flat device arenas.
static pa_status upload(pa_engine *e)
{
    pa_status st = ensure_device(e);
    if (st != PA_OK) return st;
    HIPCHK(e, hipSetDevice(e->device));
    if (!e->dev_dirty) return PA_OK;
    HIPCHK(e, hipDeviceSynchronize());   // no launch on any stream (serving lanes too) may still read the old arenas
    free_arenas(e);
    const int S = e->tm.S, N = e->tm.N;
    std::vector<ElemDev> d(e->elems.size());
    size_t nc = 0, nw = 0, ns = 0, nt = 0;
    for (size_t i = 0; i < e->elems.size(); i++) {
        const HostElem &E = e->elems[i];
        ElemDev &x = d[i];
        memset(&x, 0, sizeof x);
        if (E.filtered) continue;
        x.GG = E.GG; x.P = E.P; x.Ppad = E.Ppad; x.root = E.root;
        x.n_nodes = (int)E.nodes.size(); x.n_int = (int)E.internal.size(); x.n_up = E.n_up; x.n_down = E.n_down;
        x.has_amb = E.has_amb;
        x.off_codes = nc; nc += (size_t)S * E.Ppad;
        x.off_w = nw; nw += E.Ppad;
        x.off_site = ns; ns += (E.GG + 15) & ~15;
        x.off_tab = nt; nt += tab_words(N);
    }
    std::vector<uint8_t> codes(std::max<size_t>(nc, 1));
    std::vector<int32_t> w(std::max<size_t>(nw, 1), 0), site(std::max<size_t>(ns, 1), 0), tab(std::max<size_t>(nt, 2), 0);
    for (size_t i = 0; i < e->elems.size(); i++) {
        const HostElem &E = e->elems[i];
        if (E.filtered) continue;
        memcpy(&codes[d[i].off_codes], E.codes.data(), E.codes.size());
        memcpy(&w[d[i].off_w], E.w.data(), E.w.size() * 4);
        memcpy(&site[d[i].off_site], E.site.data(), E.site.size() * 4);
        memcpy(&tab[d[i].off_tab], E.tab.data(), E.tab.size() * 4);
    }
    HIPCHK(e, hipMalloc(&e->d_elems, sizeof(ElemDev) * std::max<size_t>(d.size(), 1)));
    HIPCHK(e, hipMalloc(&e->d_codes, codes.size()));
    HIPCHK(e, hipMalloc(&e->d_w, w.size() * 4));
    HIPCHK(e, hipMalloc(&e->d_site, site.size() * 4));
    HIPCHK(e, hipMalloc(&e->d_tab, tab.size() * 4));
    if (!d.empty()) HIPCHK(e, hipMemcpy(e->d_elems, d.data(), sizeof(ElemDev) * d.size(), hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(e->d_codes, codes.data(), codes.size(), hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(e->d_w, w.data(), w.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(e->d_site, site.data(), site.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(e->d_tab, tab.data(), tab.size() * 4, hipMemcpyHostToDevice));
    e->max_ppad = 0;
    e->mean_p = 0;
    size_t kept = 0;
    for (const HostElem &E : e->elems)
        if (!E.filtered) { e->max_ppad = std::max(e->max_ppad, E.Ppad); e->mean_p += E.P; kept++; }
    if (kept) e->mean_p /= (double)kept;
    e->dev_elems = std::move(d);
    e->dev_dirty = false;
    e->arena_gen++;
    return PA_OK;
}

namespace {

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

// Block size of a batched launch from the elements' mean pattern count.
// Default block size (no tuning): by the mean pattern count (measured on the
// 43-species workloads: 128 threads at a mean of 136 patterns, 256 at 299),
// and 256 when the batch has fewer elements than four per CU, where more
// threads per element is all the device can add.
static int block_for_mean(double meanP) { return meanP <= 96 ? 64 : (meanP <= 200 ? 128 : 256); }

static int pick_block(pa_engine *e, const int32_t *work, int32_t n)
{
    double meanP = 0;
    for (int i = 0; i < n; i++) meanP += e->elems[work[i]].P;
    meanP /= std::max(1, n);
    if (e->n_cu > 0 && n < 4 * e->n_cu) return 256;
    return block_for_mean(meanP);
}

// Workgroups resident on one CU at once, from the runtime's occupancy query
// of the kernel itself (registers under its launch bound, LDS, waves); the
// LDS and wave count alone bound it when the query fails.
static int fit_per_cu(pa_engine *e, size_t smem, int BS, int queried)
{
    const int per_cu = (int)std::min<size_t>(8, e->lds_per_cu / smem);
    const int bound = std::max(1, std::min(per_cu, 2048 / BS));
    return queried > 0 ? std::min(queried, bound) : bound;
}
static int chain_grid_max(pa_engine *e, int fp64, int BS, int shape)
{
    const int N = e->tm.N, S = e->tm.S;
    return e->n_cu * fit_per_cu(e, chain::chain_smem_bytes(N, S, BS, fp64), BS, chain::chain_blocks_per_cu(fp64, BS, N, S, shape));
}
static int loglik_grid_max(pa_engine *e, int fp64, int BS)
{
    const int N = e->tm.N;
    return e->n_cu * fit_per_cu(e, loglik_smem_bytes(N, fp64), BS, loglik_blocks_per_cu(fp64, BS, N));
}

// Scratch of one persistent launch of `grid` workgroups over elements of
// padded pattern counts ppad[] (work order).  Each workgroup's slot has a
// fixed part and a partial region for elements of up to `cap` patterns; an
// element above `cap` keeps its partials in a region of its own behind the
// slots (big_off[w], bytes from the start of that area; -1: the slot).  The
// cap minimises the total: the slots of a heavy-tailed batch are not all
// sized for its longest element.
namespace {
struct ScratchPlan {
    int cap = 64;
    size_t slot = 0, slots = 0, big = 0, total = 0;
    std::vector<int64_t> big_off;
};
}  // namespace

template <typename PartBytes>
static ScratchPlan plan_scratch(const std::vector<int> &ppad, int grid, size_t fixed, PartBytes part)
{
    std::vector<int> v(ppad);
    std::sort(v.begin(), v.end(), std::greater<int>());
    ScratchPlan best;
    best.total = SIZE_MAX;
    // candidate caps: every distinct padded count; above-cap elements are a prefix of v
    size_t big = 0;
    for (size_t j = 0; j < v.size(); j++) {
        if (j > 0 && v[j] == v[j - 1]) { big += part(v[j]); continue; }
        const size_t slot = (fixed + part(v[j]) + 255) & ~(size_t)255;
        const size_t tot = slot * (size_t)grid + big;
        if (tot < best.total) { best.cap = v[j]; best.slot = slot; best.slots = slot * (size_t)grid; best.big = big; best.total = tot; }
        big += part(v[j]);
    }
    if (v.empty()) {
        best.slot = (fixed + part(64) + 255) & ~(size_t)255;
        best.slots = best.total = best.slot * (size_t)grid;
    }
    best.big_off.assign(ppad.size(), -1);
    size_t o = 0;
    for (size_t w = 0; w < ppad.size(); w++)
        if (ppad[w] > best.cap) { best.big_off[w] = (int64_t)o; o += part(ppad[w]); }
    return best;
}

// Retained likelihood scratch above this size is released when the call ends.
static const size_t PA_RETAIN_MAX = (size_t)8 << 30;

static void release_ll_buffers(pa_engine *e)
{
    (void)hipFree(e->d_ll_scratch);
    (void)hipFree(e->d_ll_len);
    (void)hipHostFree(e->h_ll_len);
    (void)hipHostFree(e->h_tab_stage);
    e->d_ll_scratch = e->d_ll_len = nullptr;
    e->h_ll_len = e->h_tab_stage = nullptr;
    e->ll_scratch_bytes = e->ll_len_bytes = e->h_ll_len_bytes = e->h_tab_stage_bytes = 0;
}

// Per-call lengths to the device: copied into a retained pinned buffer by up
// to 16 threads, then one asynchronous copy on the engine stream (the
// launch that reads them follows on that stream and the call waits for it).
static pa_status upload_lengths(pa_engine *e, const double *lengths, size_t bytes)
{
    // no copy of an earlier call may still read the staging buffer
    HIPCHK(e, hipStreamSynchronize(e->stream));
    if (bytes > e->h_ll_len_bytes) {
        (void)hipHostFree(e->h_ll_len);
        e->h_ll_len = nullptr;
        e->h_ll_len_bytes = 0;
        HIPCHK(e, hipHostMalloc(&e->h_ll_len, bytes));
        e->h_ll_len_bytes = bytes;
    }
    const size_t chunk = (size_t)4 << 20;
    const int parts = (int)std::min<size_t>(16, (bytes + chunk - 1) / chunk);
    if (parts <= 1) {
        memcpy(e->h_ll_len, lengths, bytes);
    } else {
        const size_t per = (bytes / parts + 63) & ~(size_t)63;
        std::vector<std::thread> th;
        for (int t = 0; t < parts; t++) {
            const size_t a = std::min(bytes, per * t), b = std::min(bytes, per * (t + 1));
            if (b > a) th.emplace_back([=]() { memcpy((char *)e->h_ll_len + a, (const char *)lengths + a, b - a); });
        }
        for (auto &x : th) x.join();
    }
    HIPCHK(e, hipMemcpyAsync(e->d_ll_len, e->h_ll_len, bytes, hipMemcpyHostToDevice, e->stream));
    return PA_OK;
}

// Fixed-(Z, r) log-likelihood launch: one persistent launch over ids (in
// schedule order, largest first), one scratch slot per workgroup.
// Block size of a likelihood launch.  The kernel keeps up to 168 VGPRs (3
// waves per SIMD), so a wider workgroup mostly buys lanes for long elements:
// on 20,000 120-leaf gene trees (~360 patterns) 128 threads beat 256 by 7 %
// in FP32 (8.5 vs 9.2 ms), while FP64, with twice the partial traffic per
// lane, is fastest at 256 (11.9 vs 13.5 ms).
static int loglik_block(pa_engine *e, const int32_t *work, int32_t n, int fp64)
{
    if (e->n_cu > 0 && n < 4 * e->n_cu) return 256;
    double meanP = 0;
    for (int i = 0; i < n; i++) meanP += e->elems[work[i]].P;
    meanP /= std::max(1, n);
    if (meanP <= 96) return 64;
    if (fp64) return meanP <= 200 ? 128 : 256;
    return meanP <= 600 ? 128 : 256;
}

static pa_status run_kernel(pa_engine *e, const int32_t *ids, int32_t n, int fp64, LoglikArgs &A, DevBuf *bufs)
{
    const int S = e->tm.S, N = e->tm.N;
    const int BS = loglik_block(e, ids, n, fp64);
    const size_t smem = loglik_smem_bytes(N, fp64);
    if (smem > e->lds_per_cu) return fail(e, PA_E_INPUT, "tree too large for the LDS layout");
    const int grid = std::min(n, loglik_grid_max(e, fp64, BS));
    // resident partials come with their own arena and offsets
    const bool resident = A.alt != nullptr;
    std::vector<int> ppad(resident ? 0 : n);
    for (int i = 0; i < (int)ppad.size(); i++) ppad[i] = e->elems[ids[i]].Ppad;
    const ScratchPlan pl = resident ? ScratchPlan()
                                    : plan_scratch(ppad, grid, 0, [&](int p) { return loglik_slot_bytes(N, S, p, fp64); });
    // the previous call synchronised its stream, so its scratch is free to reuse
    if (!resident && pl.total > e->ll_scratch_bytes) {
        (void)hipFree(e->d_ll_scratch);
        e->d_ll_scratch = nullptr;
        e->ll_scratch_bytes = 0;
        HIPCHK(e, hipMalloc(&e->d_ll_scratch, pl.total));
        e->ll_scratch_bytes = pl.total;
    }
    HIPCHK(e, hipMalloc(&bufs[1].p, sizeof(int32_t) * n + sizeof(int64_t) * n + 32));
    int64_t *d_bo = (int64_t *)((char *)bufs[1].p + ((sizeof(int32_t) * n + 15) & ~(size_t)15));
    unsigned long long *d_cnt = (unsigned long long *)(d_bo + n);
    HIPCHK(e, hipMemcpyAsync(bufs[1].p, ids, sizeof(int32_t) * n, hipMemcpyHostToDevice, e->stream));
    if (!resident) HIPCHK(e, hipMemcpyAsync(d_bo, pl.big_off.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemsetAsync(d_cnt, 0, 8, e->stream));
    A.counters = d_cnt;
    A.elems = e->d_elems; A.codes = e->d_codes; A.wts = e->d_w; A.tab = e->d_tab;
    A.n_elems = (int)e->elems.size();
    A.work = (const int32_t *)bufs[1].p;
    A.n_work = n;
    A.queue = e->d_queue;
    if (!resident) {
        A.scratch = (char *)e->d_ll_scratch;
        A.slot_bytes = (int64_t)pl.slot;
        A.big = (char *)e->d_ll_scratch + pl.slots;
        A.big_off = pl.big ? d_bo : nullptr;
    }
    HIPCHK(e, launch_reset(A.queue, e->stream));
    HIPCHK(e, hipEventRecord(e->ll_t0, e->stream));
    HIPCHK(e, launch_loglik(e->M, A, fp64, BS, grid, e->stream));
    HIPCHK(e, hipEventRecord(e->ll_t1, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->ll_last_plan_bytes = pl.total;
    float ms = 0;
    unsigned long long cnt = 0;
    HIPCHK(e, hipEventElapsedTime(&ms, e->ll_t0, e->ll_t1));
    HIPCHK(e, hipMemcpy(&cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost));
    e->ll_stats.elapsed_ms = ms;
    e->ll_stats.algorithmic_bytes = (double)cnt;
    e->ll_stats.grid = grid;
    e->ll_stats.block_threads = BS;
    e->ll_stats.lds_bytes = (int32_t)smem;
    e->ll_stats.n_launches = 1;
    e->ll_stats.scratch_bytes = (double)pl.total;
    return PA_OK;
}

// Chain launch (pa_chain.hip).  Batched: one persistent launch, block size
// from the mean pattern count, workgroups pulling elements (largest first)
// from the work counter.  Per element: one single-workgroup launch per
// element, the plain per-element baseline.
static pa_status run_chain_kernel(pa_engine *e, const int32_t *ids, int32_t n, int fp64, int mode, int bs_req,
                                  chain::ChainArgs &A, DevBuf *bufs, pa_run_stats *stats)
{
    const int S = e->tm.S, N = e->tm.N, T = e->M.T;
    int BS = bs_req ? bs_req : (e->tune.block_threads ? e->tune.block_threads : pick_block(e, ids, n));
    if (BS != 64 && BS != 128 && BS != 256) return fail(e, PA_E_INPUT, "block_threads must be 64, 128 or 256");
    const size_t smem = chain::chain_smem_bytes(N, S, BS, fp64);
    if (smem > e->lds_per_cu) return fail(e, PA_E_INPUT, "tree too large for the LDS layout");
    const int shape = e->tune.shape;
    const int grid = mode == PA_RUN_PER_ELEMENT ? 1 : std::min(n, chain_grid_max(e, fp64, BS, shape));
    std::vector<int> ppad(n);
    for (int i = 0; i < n; i++) ppad[i] = e->elems[ids[i]].Ppad;
    size_t off[4];
    const size_t fixed = chain::slot_scratch_bytes(N, S, 0, fp64, off);
    const ScratchPlan pl = plan_scratch(ppad, grid, fixed, [&](int p) { return chain::chain_part_bytes(N, S, p, fp64); });
    chain::slot_scratch_bytes(N, S, pl.cap, fp64, off);
    HIPCHK(e, hipMalloc(&bufs[0].p, pl.total));
    const size_t ids_b = (sizeof(int32_t) * n + 15) & ~(size_t)15;
    HIPCHK(e, hipMalloc(&bufs[1].p, ids_b + sizeof(int64_t) * n + 64));
    int64_t *d_bo = (int64_t *)((char *)bufs[1].p + ids_b);
    unsigned long long *d_cnt = (unsigned long long *)(d_bo + n);
    HIPCHK(e, hipMemsetAsync(d_cnt, 0, 16, e->stream));
    HIPCHK(e, hipMemcpyAsync(bufs[1].p, ids, sizeof(int32_t) * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemcpyAsync(d_bo, pl.big_off.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, e->stream));
    A.counters = d_cnt;
    A.elems = e->d_elems; A.codes = e->d_codes; A.wts = e->d_w; A.site = e->d_site; A.tab = e->d_tab;
    A.queue = e->d_queue;
    A.scratch = (char *)bufs[0].p;
    A.slot_bytes = (int64_t)pl.slot;
    A.off_plin = off[0]; A.off_logt = off[1]; A.off_amb = off[2]; A.off_parts = off[3];
    A.big = (char *)bufs[0].p + pl.slots;
    A.big_off = pl.big ? d_bo : nullptr;
    hipEvent_t t0, t1;
    HIPCHK(e, hipEventCreate(&t0));
    HIPCHK(e, hipEventCreate(&t1));
    HIPCHK(e, hipEventRecord(t0, e->stream));
    if (mode == PA_RUN_PER_ELEMENT) {
        for (int i = 0; i < n; i++) {
            chain::ChainArgs B = A;
            B.work = (const int32_t *)bufs[1].p + i;
            B.n_work = 1;
            B.big_off = A.big_off ? A.big_off + i : nullptr;
            // per-element outputs shift with the work position
            B.res_d = A.res_d ? A.res_d + (size_t)i * 3 * chain::RES_D : nullptr;
            B.res_i = A.res_i ? A.res_i + (size_t)i * 3 * chain::RES_I : nullptr;
            B.maxZ = A.maxZ ? A.maxZ + (size_t)i * 3 * N : nullptr;
            B.tr_d = A.tr_d ? A.tr_d + (size_t)i * 3 * 7 * T : nullptr;
            B.tr_Z = A.tr_Z ? A.tr_Z + (size_t)i * 3 * T * N : nullptr;
            HIPCHK(e, launch_reset(e->d_queue, e->stream));
            HIPCHK(e, chain::launch_chain(e->M, B, fp64, BS, 1, e->stream, shape));
        }
    } else {
        A.work = (const int32_t *)bufs[1].p;
        A.n_work = n;
        HIPCHK(e, launch_reset(e->d_queue, e->stream));
        HIPCHK(e, chain::launch_chain(e->M, A, fp64, BS, grid, e->stream, shape));
    }
    HIPCHK(e, hipEventRecord(t1, e->stream));
    HIPCHK(e, hipEventSynchronize(t1));
    float ms = 0;
    HIPCHK(e, hipEventElapsedTime(&ms, t0, t1));
    if (stats) {
        unsigned long long cnt = 0;
        HIPCHK(e, hipMemcpy(&cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost));
        stats->elapsed_ms = ms;
        stats->algorithmic_bytes = (double)cnt;
        stats->grid = grid;
        stats->block_threads = BS;
        stats->lds_bytes = (int32_t)smem;
        stats->n_launches = mode == PA_RUN_PER_ELEMENT ? n : 1;
        stats->scratch_bytes = (double)pl.total;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return PA_OK;
}

static pa_status check_ids(pa_engine *e, const int32_t *ids, int32_t n)
{
    for (int i = 0; i < n; i++) {
        if (ids[i] < 0 || ids[i] >= (int)e->elems.size()) return fail(e, PA_E_INPUT, "element id out of range");
        if (e->elems[ids[i]].filtered) return fail(e, PA_E_INPUT, "element " + std::to_string(ids[i]) + " is filtered");
    }
    return PA_OK;
}

static std::vector<int32_t> schedule(pa_engine *e, const int32_t *ids, int32_t n, int mode);

static std::vector<int32_t> schedule(pa_engine *e, const int32_t *ids, int32_t n, int mode)
{
    std::vector<int32_t> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    if (mode != PA_RUN_BATCHED || n == 0) return order;
    {
        std::lock_guard<std::mutex> lock(e->sched_mu);
        if ((int)e->sched_ids.size() == n && memcmp(e->sched_ids.data(), ids, sizeof(int32_t) * n) == 0)
            return e->sched_order;
    }
    // longest first (GG * 3 + P descending), ties in batch order: one sort of
    // packed (inverted key, position) words
    std::vector<uint64_t> kk(n);
    for (int i = 0; i < n; i++) {
        const HostElem &x = e->elems[ids[i]];
        const int64_t key = (int64_t)x.GG * 3 + x.P;
        kk[i] = ((uint64_t)(INT64_C(0x7fffffff) - std::min<int64_t>(key, INT64_C(0x7fffffff))) << 32) | (uint32_t)i;
    }
    std::sort(kk.begin(), kk.end());
    for (int i = 0; i < n; i++) order[i] = (int32_t)(kk[i] & 0xffffffffu);
    std::lock_guard<std::mutex> lock(e->sched_mu);
    e->sched_ids.assign(ids, ids + n);
    e->sched_order = order;
    return order;
}

extern "C" pa_status pa_loglik(pa_engine *e, const int32_t *ids, int32_t n, const int32_t *Z, const double *rates,
                               int32_t precision, double *out)
{
    if (!e || !ids || !Z || !rates || !out || n < 0) return fail(e, PA_E_INPUT, "null argument");
    if (n == 0) return PA_OK;
    pa_status st = check_ids(e, ids, n);
    if (st != PA_OK) return st;
    if ((st = upload(e)) != PA_OK) return st;
    HIPCHK(e, hipSetDevice(e->device));
    const int N = e->tm.N;
    std::vector<int32_t> order = schedule(e, ids, n, PA_RUN_BATCHED);
    std::vector<int32_t> Zw((size_t)n * N);
    std::vector<double> rw((size_t)n * 2);
    for (int i = 0; i < n; i++) {
        memcpy(&Zw[(size_t)i * N], Z + (size_t)order[i] * N, sizeof(int32_t) * N);
        rw[2 * i] = rates[2 * order[i]];
        rw[2 * i + 1] = rates[2 * order[i] + 1];
    }
    DevBuf bufs[6];
    HIPCHK(e, hipMalloc(&bufs[3].p, Zw.size() * 4));
    HIPCHK(e, hipMalloc(&bufs[4].p, rw.size() * 8));
    HIPCHK(e, hipMalloc(&bufs[5].p, (size_t)n * 8));
    HIPCHK(e, hipMemcpy(bufs[3].p, Zw.data(), Zw.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(bufs[4].p, rw.data(), rw.size() * 8, hipMemcpyHostToDevice));
    LoglikArgs A;
    memset(&A, 0, sizeof A);
    A.Zin = (const int32_t *)bufs[3].p;
    A.rates_in = (const double *)bufs[4].p;
    A.ll_out = (double *)bufs[5].p;
    std::vector<int32_t> sorted(n);
    for (int i = 0; i < n; i++) sorted[i] = ids[order[i]];
    st = run_kernel(e, sorted.data(), n, precision == PA_FP64, A, bufs);
    if (st != PA_OK) return st;
    std::vector<double> ll(n);
    HIPCHK(e, hipMemcpy(ll.data(), bufs[5].p, (size_t)n * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) out[order[i]] = ll[i];
    if (e->ll_scratch_bytes + e->ll_len_bytes > PA_RETAIN_MAX) release_ll_buffers(e);
    return PA_OK;
}

// Log-likelihood under the gBGC substitution model (PhyloAcc-ST-GBGC): every
// branch's P(t) comes from Q(s, b) of its class -- background (Z = 0),
// conserved (Z = 1) or accelerated (Z = 2) selection, with or without the
// GC-biased conversion bias when B = 1 -- over the element's own branch
// lengths (BPP_C::getRate / getTransition / init_cache_TM,
// src/PhyloAcc-ST-GBGC/SRC/bpp_c.cpp:13-29,1405-1427,1563-1621; the rate
// matrices of bpp_c.cpp:191-197).  The six class eigensystems of each element
// are built on the device, then the likelihood kernel prunes as pa_loglik.
extern "C" pa_status pa_loglik_gbgc(pa_engine *e, const int32_t *ids, int32_t n, const int32_t *Z, const int32_t *B,
                                    const double *sel, int32_t precision, double *out)
{
    if (!e || !ids || !Z || !B || !sel || !out || n < 0) return fail(e, PA_E_INPUT, "null argument");
    if (n == 0) return PA_OK;
    pa_status st = check_ids(e, ids, n);
    if (st != PA_OK) return st;
    if ((st = upload(e)) != PA_OK) return st;
    HIPCHK(e, hipSetDevice(e->device));
    const int N = e->tm.N;
    std::vector<int32_t> order = schedule(e, ids, n, PA_RUN_BATCHED);
    std::vector<int32_t> Zw((size_t)n * N), Bw((size_t)n * N);
    std::vector<double> sw((size_t)n * 3);
    for (int i = 0; i < n; i++) {
        memcpy(&Zw[(size_t)i * N], Z + (size_t)order[i] * N, sizeof(int32_t) * N);
        memcpy(&Bw[(size_t)i * N], B + (size_t)order[i] * N, sizeof(int32_t) * N);
        memcpy(&sw[(size_t)i * 3], sel + (size_t)order[i] * 3, sizeof(double) * 3);
    }
    DevBuf bufs[8];
    HIPCHK(e, hipMalloc(&bufs[3].p, Zw.size() * 4));
    HIPCHK(e, hipMalloc(&bufs[4].p, sw.size() * 8));
    HIPCHK(e, hipMalloc(&bufs[5].p, (size_t)n * 8));
    HIPCHK(e, hipMalloc(&bufs[6].p, Bw.size() * 4));
    HIPCHK(e, hipMalloc(&bufs[7].p, (size_t)n * GB_CLASSES * GB_EIG_WORDS * 8));
    HIPCHK(e, hipMemcpy(bufs[3].p, Zw.data(), Zw.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(bufs[4].p, sw.data(), sw.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(bufs[6].p, Bw.data(), Bw.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(e, launch_gbgc_eig(e->Q0, (const double *)bufs[4].p, n, (double *)bufs[7].p, e->stream));
    LoglikArgs A;
    memset(&A, 0, sizeof A);
    A.Zin = (const int32_t *)bufs[3].p;
    A.Bin = (const int32_t *)bufs[6].p;
    A.gb_eig = (const double *)bufs[7].p;
    A.ll_out = (double *)bufs[5].p;
    std::vector<int32_t> sorted(n);
    for (int i = 0; i < n; i++) sorted[i] = ids[order[i]];
    st = run_kernel(e, sorted.data(), n, precision == PA_FP64, A, bufs);
    if (st != PA_OK) return st;
    std::vector<double> ll(n);
    HIPCHK(e, hipMemcpy(ll.data(), bufs[5].p, (size_t)n * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) out[order[i]] = ll[i];
    if (e->ll_scratch_bytes + e->ll_len_bytes > PA_RETAIN_MAX) release_ll_buffers(e);
    return PA_OK;
}

// Launch statistics of the last pa_loglik / pa_loglik_lengths /
// pa_loglik_gbgc call: the likelihood kernel's HIP-event duration and the
// partial, code and weight bytes its passes had to move (counted on device).
extern "C" pa_status pa_loglik_last_stats(pa_engine *e, pa_run_stats *out)
{
    if (!e || !out) return fail(e, PA_E_INPUT, "null argument");
    *out = e->ll_stats;
    return PA_OK;
}

// Log-likelihood with per-call branch lengths (every branch at rate 1):
// the pruning of an element's own tree, e.g. a gene tree whose lengths come
// from pa_gene_lengths.
// Own-tree log-likelihoods at per-call lengths, with the model's base
// frequencies or, when piA is given, each element's strand-symmetric ones.
// BPP's exchangeabilities from RATE_MAT and BACKGROUND, with the reference's
// assignment (src/PhyloAcc-GT/bpp.hpp:346-353: CG takes AC's value, CT takes
// C->G / pi_G, GT takes AG's value).
static void gt_exchangeabilities(const pa_engine *e, double *inst)
{
    const double *Q = e->Q0, *p0 = e->M.pi;
    inst[0] = Q[0 * 4 + 1] / p0[1];
    inst[1] = Q[0 * 4 + 2] / p0[2];
    inst[2] = Q[0 * 4 + 3] / p0[3];
    inst[3] = inst[0];
    inst[4] = Q[1 * 4 + 2] / p0[2];
    inst[5] = inst[1];
}

static pa_status loglik_lengths_impl(pa_engine *e, const int32_t *ids, int32_t n, const double *lengths,
                                     const double *piA, int32_t precision, double *out,
                                     const int32_t *lens0 = nullptr, const int32_t *lens1 = nullptr,
                                     double *out1 = nullptr, const GtExtra *gx = nullptr)
{
    const bool dev_len = gx && gx->d_len;
    if (!e || !ids || (!lengths && !dev_len) || !out || n < 0) return fail(e, PA_E_INPUT, "null argument");
    if (n == 0) return PA_OK;
    pa_status st = check_ids(e, ids, n);
    if (st != PA_OK) return st;
    if ((st = upload(e)) != PA_OK) return st;
    const int N = e->tm.N;
    std::vector<int32_t> order = schedule(e, ids, n, PA_RUN_BATCHED);
    // the lengths go up in the caller's row order; work position w reads row order[w]
    const size_t lbytes = (size_t)n * N * 8;
    DevBuf bufs[10];
    HIPCHK(e, hipMalloc(&bufs[5].p, (size_t)n * 8));
    HIPCHK(e, hipMalloc(&bufs[6].p, (size_t)n * 4));
    if (!dev_len && lbytes > e->ll_len_bytes) {
        (void)hipFree(e->d_ll_len);
        e->d_ll_len = nullptr;
        e->ll_len_bytes = 0;
        HIPCHK(e, hipMalloc(&e->d_ll_len, lbytes));
        e->ll_len_bytes = lbytes;
    }
    if (!dev_len && (st = upload_lengths(e, lengths, lbytes)) != PA_OK) return st;
    HIPCHK(e, hipMemcpy(bufs[6].p, order.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    LoglikArgs A;
    memset(&A, 0, sizeof A);
    A.ll_out = (double *)bufs[5].p;
    A.len_in = dev_len ? gx->d_len : (const double *)e->d_ll_len;
    if (gx) { A.row_out0 = gx->d_row0; A.row_out1 = gx->d_row1; }
    A.len_row = (const int32_t *)bufs[6].p;
    if (piA) {
        double inst[6];
        gt_exchangeabilities(e, inst);
        std::vector<double> pw(n);
        for (int i = 0; i < n; i++) pw[i] = piA[order[i]];
        HIPCHK(e, hipMalloc(&bufs[7].p, (size_t)n * 8));
        HIPCHK(e, hipMalloc(&bufs[8].p, (size_t)n * GB_EIG_WORDS * 8));
        HIPCHK(e, hipMalloc(&bufs[9].p, (size_t)n * 4 * 8));
        HIPCHK(e, hipMemcpy(bufs[7].p, pw.data(), (size_t)n * 8, hipMemcpyHostToDevice));
        HIPCHK(e, launch_gt_eig(inst, (const double *)bufs[7].p, n, (double *)bufs[8].p, (double *)bufs[9].p, e->stream));
        A.el_eig = (const double *)bufs[8].p;
        A.el_pi = (const double *)bufs[9].p;
    }
    std::vector<int32_t> sorted(n);
    for (int i = 0; i < n; i++) sorted[i] = ids[order[i]];
    DevBuf pre[2];
    if (lens0) {
        // the two site prefixes per caller row; the root sums walk the sites
        std::vector<int32_t> pl((size_t)n * 2);
        for (int i = 0; i < n; i++) { pl[2 * i] = lens0[i]; pl[2 * i + 1] = lens1[i]; }
        HIPCHK(e, hipMalloc(&pre[0].p, pl.size() * 4));
        HIPCHK(e, hipMalloc(&pre[1].p, (size_t)n * 8));
        HIPCHK(e, hipMemcpy(pre[0].p, pl.data(), pl.size() * 4, hipMemcpyHostToDevice));
        A.pre_len = (const int32_t *)pre[0].p;
        A.ll_out2 = (double *)pre[1].p;
        A.site = e->d_site;
    }
    // resident partials: every work position's region in the caller's arena
    DevBuf rb;
    if (gx && gx->d_big) {
        if (!gx->d_alt || !gx->row_big_off) return fail(e, PA_E_INPUT, "resident partials need slot bits and offsets");
        std::vector<int64_t> bo(n);
        for (int i = 0; i < n; i++) bo[i] = gx->row_big_off[order[i]];
        HIPCHK(e, hipMalloc(&rb.p, (size_t)n * 8));
        HIPCHK(e, hipMemcpy(rb.p, bo.data(), (size_t)n * 8, hipMemcpyHostToDevice));
        A.big = gx->d_big;
        A.big_off = (const int64_t *)rb.p;
        A.alt = gx->d_alt;
        A.path_node = gx->d_path;
    }
    // gene-tree moves: per-row trees, history draws, candidate scores (rows in the caller's order)
    DevBuf gb[12];
    auto up = [&](DevBuf &b, const void *src, size_t bytes) -> hipError_t {
        hipError_t r = hipMalloc(&b.p, std::max<size_t>(bytes, 8));
        if (r == hipSuccess && bytes) r = hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice);
        return r;
    };
    if (gx) {
        A.site = e->d_site;
        if (gx->tabs) {
            HIPCHK(e, up(gb[0], gx->tabs, (size_t)n * tab_words(N) * 4));
            HIPCHK(e, up(gb[1], gx->hdrs, (size_t)n * 4 * 4));
            A.tab_call = (const int32_t *)gb[0].p;
            A.hdr_call = (const int32_t *)gb[1].p;
        }
        if (gx->d_hist) {
            HIPCHK(e, up(gb[2], gx->hist_off, (size_t)n * 8));
            HIPCHK(e, up(gb[3], gx->hist_ns, (size_t)n * 4));
            A.hist = gx->d_hist;
            A.hist_off = (const int64_t *)gb[2].p;
            A.hist_ns = (const int32_t *)gb[3].p;
            if (gx->hist_draw) {
                HIPCHK(e, up(gb[4], gx->hist_len, (size_t)n * 4));
                A.hist_len = (const int32_t *)gb[4].p;
                A.hist_draw = 1;
                A.hist_sweep = gx->hist_sweep;
            }
        }
        if (gx->cand_score) {
            const size_t rows = (size_t)n * gx->cand_max;
            HIPCHK(e, up(gb[5], gx->cand_branch, (size_t)n * 4));
            HIPCHK(e, up(gb[6], gx->cand_n, (size_t)n * 4));
            HIPCHK(e, up(gb[7], gx->cand_node, rows * 4));
            HIPCHK(e, up(gb[8], gx->cand_kind, rows * 4));
            HIPCHK(e, up(gb[9], gx->cand_sites, (size_t)n * 4));
            HIPCHK(e, up(gb[10], gx->cand_len, rows * 8));
            HIPCHK(e, hipMalloc(&gb[11].p, std::max<size_t>(rows * 8, 8)));
            A.cand_branch = (const int32_t *)gb[5].p;
            A.cand_n = (const int32_t *)gb[6].p;
            A.cand_node = (const int32_t *)gb[7].p;
            A.cand_kind = (const int32_t *)gb[8].p;
            A.cand_sites = (const int32_t *)gb[9].p;
            A.cand_len = (const double *)gb[10].p;
            A.cand_max = gx->cand_max;
            A.cand_score = (double *)gb[11].p;
        }
    }
    st = run_kernel(e, sorted.data(), n, precision == PA_FP64, A, bufs);
    if (st != PA_OK) return st;
    std::vector<double> ll(n);
    HIPCHK(e, hipMemcpy(ll.data(), bufs[5].p, (size_t)n * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) out[order[i]] = ll[i];
    if (lens0) {
        HIPCHK(e, hipMemcpy(ll.data(), pre[1].p, (size_t)n * 8, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; i++) out1[order[i]] = ll[i];
    }
    if (gx && gx->cand_score)
        HIPCHK(e, hipMemcpy(gx->cand_score, gb[11].p, (size_t)n * gx->cand_max * 8, hipMemcpyDeviceToHost));
    if (e->ll_scratch_bytes + e->ll_len_bytes > PA_RETAIN_MAX) release_ll_buffers(e);
    return PA_OK;
}

// Element id's node table under the tree children [N*2] (the root is the
// node no other node has as a child), as a gene tree: missing leaves from
// the element, an internal node missing when both children are.
static pa_status gene_elem(pa_engine *e, int32_t id, const int32_t *children, int32_t *T, int32_t *hdr)
{
    uint8_t lm[MAXN];
    pa_status st = element_gene_tree(e, id, lm, nullptr);
    if (st != PA_OK) return st;
    const int N = e->tm.N, S = e->tm.S;
    thread_local std::vector<int> ch2, par2, post;
    thread_local std::vector<uint8_t> seen, missing;
    thread_local std::vector<double> zero;
    thread_local std::vector<std::pair<int, int>> stk;
    ch2.assign(children, children + 2 * N);
    par2.assign(N, N);
    seen.assign(N, 0);
    for (int v = 0; v < S; v++)
        if (ch2[2 * v] != -1 || ch2[2 * v + 1] != -1) return fail(e, PA_E_INPUT, "gene tree leaves must have no children");
    for (int v = S; v < N; v++)
        for (int k = 0; k < 2; k++) {
            const int x = ch2[2 * v + k];
            if (x < 0 || x >= N || x == v || seen[x]) return fail(e, PA_E_INPUT, "gene tree is not a binary tree over the leaves");
            seen[x] = 1;
            par2[x] = v;
        }
    int root = -1;
    for (int v = S; v < N; v++)
        if (!seen[v]) { if (root >= 0) return fail(e, PA_E_INPUT, "gene tree has two roots"); root = v; }
    if (root < 0) return fail(e, PA_E_INPUT, "gene tree has no root");
    post.clear();
    {
        stk.assign(1, {root, 0});
        while (!stk.empty() && (int)post.size() <= N) {
            const int v = stk.back().first, k = stk.back().second;
            if (v >= S && k < 2) { stk.back().second++; stk.push_back({ch2[2 * v + k], 0}); continue; }
            post.push_back(v);
            stk.pop_back();
        }
    }
    if ((int)post.size() != N) return fail(e, PA_E_INPUT, "gene tree does not reach every node");
    missing.assign(N, 0);
    for (int v : post) missing[v] = v < S ? lm[v] : (missing[ch2[2 * v]] && missing[ch2[2 * v + 1]]);
    zero.assign(N, 0.0);
    const std::vector<int> *nodes, *internal;
    int n_up = 0, n_down = 0;
    fill_table(e->tm, ch2, par2, zero.data(), missing, true, root, T, n_up, n_down, &nodes, &internal);
    hdr[0] = root; hdr[1] = (int32_t)nodes->size(); hdr[2] = n_up; hdr[3] = n_down;
    return PA_OK;
}

namespace pa {

pa_status gene_table(pa_engine *e, int32_t id, const int32_t *children, int32_t *T, int32_t *hdr)
{
    return gene_elem(e, id, children, T, hdr);
}

// Many elements' trees at once, from their node tables (gene_table's):
// the host copies change, and when the device copy is current one upload of
// the new tables and records and one scatter launch replace them there.
pa_status element_set_trees(pa_engine *e, const int32_t *ids, int n, const int32_t *const *tabs,
                            const int32_t *const *hdrs)
{
    if (n <= 0) return PA_OK;
    const int N = e->tm.N, S = e->tm.S, W = tab_words(N), D = tab_dist(N);
    for (int k = 0; k < n; k++)
        if (ids[k] < 0 || ids[k] >= (int)e->elems.size() || hdrs[k][1] < 1 || hdrs[k][1] > N)
            return fail(e, PA_E_INPUT, "bad element or node table");
    const bool dev = !e->dev_dirty && e->d_tab && e->d_elems;
    // the rows that go to the device (elements with a device record), in
    // batch order, staged in a retained pinned buffer
    std::vector<int> row(n, -1);
    int ns = 0;
    if (dev)
        for (int k = 0; k < n; k++)
            if (ids[k] < (int)e->dev_elems.size()) row[k] = ns++;
    const size_t stage_bytes = (size_t)ns * W * 4;
    if (stage_bytes > e->h_tab_stage_bytes) {
        (void)hipHostFree(e->h_tab_stage);
        e->h_tab_stage = nullptr;
        e->h_tab_stage_bytes = 0;
        HIPCHK(e, hipHostMalloc(&e->h_tab_stage, stage_bytes));
        e->h_tab_stage_bytes = stage_bytes;
    }
    int32_t *stage = (int32_t *)e->h_tab_stage;
    std::vector<int32_t> dst_id(ns);
    std::vector<int64_t> dst_off(ns);
    std::vector<ElemDev> recs(ns);
    // one element per index on each thread (the ids are distinct, so no two
    // threads touch one element)
    auto one = [&](int k) {
        const int id = ids[k];
        HostElem &E = e->elems[id];
        const int32_t *T = tabs[k], *hd = hdrs[k];
        // gene trees take their branch lengths per call; the stored ones stay
        if ((int)E.tab.size() == W) {
            memcpy(E.tab.data(), T, (size_t)D * 4);
        } else {
            std::vector<int32_t> tab(T, T + W);
            if ((int)E.tab.size() >= W) memcpy(tab.data() + D, E.tab.data() + D, (size_t)(W - D) * 4);
            E.tab = std::move(tab);
        }
        E.root = hd[0];
        E.n_up = hd[2];
        E.n_down = hd[3];
        compact_kids(E.tab.data() + tab_ch0(N), E.tab.data() + tab_ch1(N), S, N, E.kids);
        E.nodes.assign(T + tab_nodes(N), T + tab_nodes(N) + hd[1]);
        E.internal.assign(T + tab_int(N), T + tab_int(N) + (hd[1] - 1) / 2);
        const int r = row[k];
        if (r < 0) return;
        ElemDev &x = e->dev_elems[id];
        x.root = E.root; x.n_nodes = (int)E.nodes.size(); x.n_int = (int)E.internal.size();
        x.n_up = E.n_up; x.n_down = E.n_down;
        memcpy(&stage[(size_t)r * W], E.tab.data(), (size_t)W * 4);
        dst_id[r] = id;
        dst_off[r] = x.off_tab;
        recs[r] = x;
    };
    const int nt = n < 256 ? 1 : (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (nt == 1) {
        for (int k = 0; k < n; k++) one(k);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; t++)
            th.emplace_back([&, t]() {
                for (int k = t; k < n; k += nt) one(k);
            });
        for (auto &x : th) x.join();
    }
    if (ns == 0) return PA_OK;
    DevBuf b[4];
    HIPCHK(e, hipMalloc(&b[0].p, stage_bytes));
    HIPCHK(e, hipMalloc(&b[1].p, (size_t)ns * sizeof(ElemDev)));
    HIPCHK(e, hipMalloc(&b[2].p, (size_t)ns * 8));
    HIPCHK(e, hipMalloc(&b[3].p, (size_t)ns * 4));
    HIPCHK(e, hipMemcpy(b[0].p, stage, stage_bytes, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(b[1].p, recs.data(), (size_t)ns * sizeof(ElemDev), hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(b[2].p, dst_off.data(), (size_t)ns * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(b[3].p, dst_id.data(), (size_t)ns * 4, hipMemcpyHostToDevice));
    HIPCHK(e, launch_gt_scatter_tables((const int32_t *)b[0].p, W, (const int64_t *)b[2].p, e->d_tab,
                                       (const ElemDev *)b[1].p, (const int32_t *)b[3].p, e->d_elems, ns, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return PA_OK;
}

GtHistory &engine_history(pa_engine *e) { return e->hist; }

int element_ppad(pa_engine *e, int32_t id)
{
    return id >= 0 && id < (int)e->elems.size() ? e->elems[id].Ppad : -1;
}

pa_status engine_stream(pa_engine *e, hipStream_t *st)
{
    pa_status r = ensure_device(e);
    if (r != PA_OK) return r;
    if ((r = upload(e)) != PA_OK) return r;
    *st = e->stream;
    return PA_OK;
}

pa_status engine_history_reserve(pa_engine *e, size_t bytes)
{
    pa_status st = ensure_device(e);
    if (st != PA_OK) return st;
    if (bytes <= e->hist.cap) return PA_OK;
    HIPCHK(e, hipDeviceSynchronize());
    (void)hipFree(e->hist.d);
    e->hist.d = nullptr;
    e->hist.cap = 0;
    e->hist.valid = false;
    HIPCHK(e, hipMalloc(&e->hist.d, bytes));
    e->hist.cap = bytes;
    return PA_OK;
}

pa_status gt_launch(pa_engine *e, const int32_t *ids, int32_t n, const double *lengths, const double *piA,
                    const int32_t *lens0, const int32_t *lens1, int32_t precision, double *out0, double *out1,
                    const GtExtra &x)
{
    if (n == 0) return PA_OK;
    pa_status st = check_ids(e, ids, n);
    if (st != PA_OK) return st;
    for (int i = 0; i < n; i++) {
        const int GG = e->elems[ids[i]].GG;
        if (lens0 && (lens1[i] < 0 || lens1[i] > lens0[i] || lens0[i] > GG))
            return fail(e, PA_E_INPUT, "prefix lengths must satisfy 0 <= lens1 <= lens0 <= the element's sites");
        if (x.hist_draw && (x.hist_len[i] < 0 || x.hist_len[i] > GG || x.hist_len[i] > x.hist_ns[i]))
            return fail(e, PA_E_INPUT, "history length beyond the element's sites or its row");
        if ((x.hist_draw || x.cand_score) && !x.d_hist) return fail(e, PA_E_INPUT, "no history arena");
        if (x.cand_score && (x.cand_sites[i] < 0 || x.cand_sites[i] > GG || x.cand_sites[i] > x.hist_ns[i] ||
                             x.cand_n[i] < 0 || x.cand_n[i] > x.cand_max))
            return fail(e, PA_E_INPUT, "candidate sites or count out of range");
        if (x.cand_score)
            for (int q = 0; q < x.cand_n[i]; q++) {
                const int t = x.cand_node[(size_t)i * x.cand_max + q], k = x.cand_kind[(size_t)i * x.cand_max + q];
                if (k < 0 || k > 2 || (k == 0 && (t < e->tm.S || t >= e->tm.N)) || x.cand_branch[i] < 0 ||
                    x.cand_branch[i] >= e->tm.N)
                    return fail(e, PA_E_INPUT, "bad candidate");
            }
    }
    return loglik_lengths_impl(e, ids, n, lengths, piA, precision, out0, lens0, lens1, out1, &x);
}

// A resident-partials likelihood launch prepared once for a fixed batch
// (the gene-tree height session): work order, block size, per-work-position
// arena offsets, the elements' eigensystems and the result rows stay on the
// device; a run uploads the prefix lengths and launches, without waiting.
struct GtPlan {
    int n = 0, fp64 = 0, BS = 0, grid = 0;
    std::vector<int32_t> pl;          // prefix lengths staged for the last run
    std::vector<int32_t> GG;          // sites per row
    DevBuf work, row, bo, pre, out, out2, eig, pi, cnt;
    LoglikArgs A;
};

pa_status gt_plan_build(pa_engine *e, const int32_t *ids, int32_t n, const double *piA, int32_t precision,
                        const GtExtra &x, GtPlan **out)
{
    *out = nullptr;
    if (n <= 0 || !x.d_len || !x.d_big || !x.row_big_off || !x.d_alt) return fail(e, PA_E_INPUT, "incomplete plan");
    pa_status st = check_ids(e, ids, n);
    if (st != PA_OK) return st;
    if ((st = upload(e)) != PA_OK) return st;
    std::unique_ptr<GtPlan> p(new GtPlan);
    const int N = e->tm.N, fp64 = precision == PA_FP64;
    p->n = n;
    p->fp64 = fp64;
    const std::vector<int32_t> order = schedule(e, ids, n, PA_RUN_BATCHED);
    std::vector<int32_t> sorted(n);
    std::vector<int64_t> bo(n);
    p->GG.resize(n);
    for (int i = 0; i < n; i++) {
        sorted[i] = ids[order[i]];
        bo[i] = x.row_big_off[order[i]];
        p->GG[i] = e->elems[ids[i]].GG;
    }
    p->BS = loglik_block(e, sorted.data(), n, fp64);
    if (loglik_smem_bytes(N, fp64) > e->lds_per_cu) return fail(e, PA_E_INPUT, "tree too large for the LDS layout");
    p->grid = std::min(n, loglik_grid_max(e, fp64, p->BS));
    HIPCHK(e, hipMalloc(&p->work.p, (size_t)n * 4));
    HIPCHK(e, hipMalloc(&p->row.p, (size_t)n * 4));
    HIPCHK(e, hipMalloc(&p->bo.p, (size_t)n * 8));
    HIPCHK(e, hipMalloc(&p->pre.p, (size_t)n * 8));
    HIPCHK(e, hipMalloc(&p->out.p, (size_t)n * 8));
    HIPCHK(e, hipMalloc(&p->out2.p, (size_t)n * 8));
    HIPCHK(e, hipMalloc(&p->cnt.p, 8));
    HIPCHK(e, hipMemcpy(p->work.p, sorted.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(p->row.p, order.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(p->bo.p, bo.data(), (size_t)n * 8, hipMemcpyHostToDevice));
    LoglikArgs &A = p->A;
    memset(&A, 0, sizeof A);
    if (piA) {
        double inst[6];
        gt_exchangeabilities(e, inst);
        std::vector<double> pw(n);
        for (int i = 0; i < n; i++) pw[i] = piA[order[i]];
        DevBuf dpw;
        HIPCHK(e, hipMalloc(&dpw.p, (size_t)n * 8));
        HIPCHK(e, hipMalloc(&p->eig.p, (size_t)n * GB_EIG_WORDS * 8));
        HIPCHK(e, hipMalloc(&p->pi.p, (size_t)n * 4 * 8));
        HIPCHK(e, hipMemcpy(dpw.p, pw.data(), (size_t)n * 8, hipMemcpyHostToDevice));
        HIPCHK(e, launch_gt_eig(inst, (const double *)dpw.p, n, (double *)p->eig.p, (double *)p->pi.p, e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
        A.el_eig = (const double *)p->eig.p;
        A.el_pi = (const double *)p->pi.p;
    }
    A.ll_out = (double *)p->out.p;
    A.len_in = x.d_len;
    A.len_row = (const int32_t *)p->row.p;
    A.row_out0 = x.d_row0;
    A.row_out1 = x.d_row1;
    A.big = x.d_big;
    A.big_off = (const int64_t *)p->bo.p;
    A.alt = x.d_alt;
    A.work = (const int32_t *)p->work.p;
    A.n_work = n;
    A.counters = (unsigned long long *)p->cnt.p;
    *out = p.release();
    return PA_OK;
}

pa_status gt_plan_run(pa_engine *e, GtPlan *p, const int32_t *lens0, const int32_t *lens1, const int32_t *d_path,
                      bool same_lens, bool add_count)
{
    pa_status st = upload(e);
    if (st != PA_OK) return st;
    const int n = p->n;
    LoglikArgs A = p->A;
    if (lens0 && same_lens && (int)p->pl.size() == 2 * n) {
        // the prefix lengths of the previous run, already on the device
        A.pre_len = (const int32_t *)p->pre.p;
        A.ll_out2 = (double *)p->out2.p;
    } else if (lens0) {
        p->pl.resize((size_t)n * 2);
        for (int i = 0; i < n; i++) {
            if (lens1[i] < 0 || lens1[i] > lens0[i] || lens0[i] > p->GG[i])
                return fail(e, PA_E_INPUT, "prefix lengths must satisfy 0 <= lens1 <= lens0 <= the element's sites");
            p->pl[2 * i] = lens0[i];
            p->pl[2 * i + 1] = lens1[i];
        }
        HIPCHK(e, hipMemcpyAsync(p->pre.p, p->pl.data(), (size_t)n * 8, hipMemcpyHostToDevice, e->stream));
        A.pre_len = (const int32_t *)p->pre.p;
        A.ll_out2 = (double *)p->out2.p;
    }
    // the engine's arenas as they are now
    A.elems = e->d_elems; A.codes = e->d_codes; A.wts = e->d_w; A.tab = e->d_tab;
    A.n_elems = (int)e->elems.size();
    A.site = e->d_site;
    A.queue = e->d_queue;
    A.path_node = d_path;
    if (!add_count) HIPCHK(e, hipMemsetAsync(p->cnt.p, 0, 8, e->stream));
    HIPCHK(e, launch_reset(A.queue, e->stream));
    HIPCHK(e, hipEventRecord(e->ll_t0, e->stream));
    HIPCHK(e, launch_loglik(e->M, A, p->fp64, p->BS, p->grid, e->stream));
    HIPCHK(e, hipEventRecord(e->ll_t1, e->stream));
    e->ll_stats.grid = p->grid;
    e->ll_stats.block_threads = p->BS;
    e->ll_stats.lds_bytes = (int32_t)loglik_smem_bytes(e->tm.N, p->fp64);
    e->ll_stats.n_launches = 1;
    e->ll_stats.scratch_bytes = 0;
    return PA_OK;
}

pa_status gt_plan_stats(pa_engine *e, GtPlan *p)
{
    HIPCHK(e, hipStreamSynchronize(e->stream));
    float ms = 0;
    unsigned long long cnt = 0;
    HIPCHK(e, hipEventElapsedTime(&ms, e->ll_t0, e->ll_t1));
    HIPCHK(e, hipMemcpy(&cnt, p->cnt.p, sizeof cnt, hipMemcpyDeviceToHost));
    e->ll_stats.elapsed_ms = ms;
    e->ll_stats.algorithmic_bytes = (double)cnt;
    return PA_OK;
}

void gt_plan_free(GtPlan *p) { delete p; }

}  // namespace pa

// The gene-tree log-likelihood over two prefixes of each element's sites,
// the first lens0[i] and the first lens1[i] (Wang-Landau blocks of
// PhyloAcc-GT's Gibbs, src/PhyloAcc-GT/bpp_c.cpp:359-380,436-443: log_lik[0]
// sums sites g < lens[0], log_lik[1] sites g < lens[1]), one launch: the
// patterns are pruned once and summed with two weight vectors.  piA may be
// null (the model's base frequencies) or per element as in
// pa_loglik_lengths_pi.
extern "C" pa_status pa_loglik_lengths_split(pa_engine *e, const int32_t *ids, int32_t n, const double *lengths,
                                             const double *piA, const int32_t *lens0, const int32_t *lens1,
                                             int32_t precision, double *out0, double *out1)
{
    if (!e || !ids || !lengths || !lens0 || !lens1 || !out0 || !out1 || n < 0) return fail(e, PA_E_INPUT, "null argument");
    pa_status st = check_ids(e, ids, n);
    if (st != PA_OK) return st;
    for (int i = 0; i < n; i++) {
        const int GG = e->elems[ids[i]].GG;
        if (lens1[i] < 0 || lens1[i] > lens0[i] || lens0[i] > GG)
            return fail(e, PA_E_INPUT, "prefix lengths must satisfy 0 <= lens1 <= lens0 <= the element's sites");
        if (piA && !(piA[i] > 0 && piA[i] < 0.5)) return fail(e, PA_E_INPUT, "piA must lie in (0, 1/2)");
    }
    return loglik_lengths_impl(e, ids, n, lengths, piA, precision, out0, lens0, lens1, out1);
}

// One Wang-Landau indicator step of PhyloAcc-GT's Gibbs (src/PhyloAcc-GT/
// bpp_c.cpp:399-440) for n elements at sweep m.  state[i*7 ..]: the two log
// normalising constants, the two momenta, the learning rate and the two
// visit counts; at m = 0 the constants start at loglik + logpz, the momenta
// at 0, the rate at 1 and the counts at 0 (bpp_c.cpp:423-427).  The
// indicator is drawn from normalise(exp(loglik + logpz - constants)) (one
// uniform, tag WL, sweep m), then momentum = 0.9 momentum - rate * prob,
// constants -= momentum and renormalised, and the rate shrinks to
// 1 / (1 + 1 / rate) (counts reset) once the visit ratio is within
// (2/3, 3/2).
extern "C" pa_status pa_wl_step(pa_engine *e, int32_t n, int32_t m, const double *loglik, const double *logpz,
                                double *state, int32_t *indicator)
{
    if (!e || n < 0 || (n > 0 && (!loglik || !logpz || !state || !indicator))) return fail(e, PA_E_INPUT, "null argument");
    const Key key{e->M.key0, e->M.key1};
    const double u = slot_uniform(key, 0, (uint32_t)m, tagword(0, TAG_WL), 0);
    for (int i = 0; i < n; i++) {
        double *s = state + (size_t)i * 7;
        const double *ll = loglik + 2 * i, *lz = logpz + 2 * i;
        if (m == 0) {
            s[0] = ll[0] + lz[0]; s[1] = ll[1] + lz[1];
            s[2] = s[3] = 0.0;
            s[4] = 1.0;
            s[5] = s[6] = 0.0;
        }
        double y[2] = {ll[0] + lz[0] - s[0], ll[1] + lz[1] - s[1]};
        const double mx = std::max(y[0], y[1]);
        const double x = mx + log(exp(y[0] - mx) + exp(y[1] - mx));
        double prob[2] = {0.0, 0.0};
        if (x != INFINITY) { prob[0] = exp(y[0] - x); prob[1] = exp(y[1] - x); }
        const int ind = categorical<2>(prob, u);
        indicator[i] = ind;
        s[2] = 0.9 * s[2] - s[4] * prob[0];
        s[3] = 0.9 * s[3] - s[4] * prob[1];
        s[5 + ind] += 1;
        s[0] -= s[2];
        s[1] -= s[3];
        const double mz = std::max(s[0], s[1]);
        const double z = mz + log(exp(s[0] - mz) + exp(s[1] - mz));
        s[0] -= z;
        s[1] -= z;
        const double rr = s[5] / s[6];
        if (rr > 2.0 / 3.0 && rr < 3.0 / 2.0) {
            s[4] = 1 / (1 + 1 / s[4]);
            s[5] = s[6] = 0.0;
        }
    }
    return PA_OK;
}

extern "C" pa_status pa_loglik_lengths(pa_engine *e, const int32_t *ids, int32_t n, const double *lengths,
                                       int32_t precision, double *out)
{
    return loglik_lengths_impl(e, ids, n, lengths, nullptr, precision, out);
}

// PhyloAcc-GT's model with each element's own strand-symmetric base
// frequencies pi = (a, 1/2 - a, 1/2 - a, a) (BPP_C::sample_pi and getQmat,
// src/PhyloAcc-GT/bpp_c.cpp:1545-1548,160-185): the element's Q is rebuilt from
// the exchangeabilities on the device, then pruned as pa_loglik_lengths.
extern "C" pa_status pa_loglik_lengths_pi(pa_engine *e, const int32_t *ids, int32_t n, const double *lengths,
                                          const double *piA, int32_t precision, double *out)
{
    if (!piA && n > 0) return fail(e, PA_E_INPUT, "null argument");
    for (int i = 0; i < n; i++)
        if (!(piA[i] > 0 && piA[i] < 0.5)) return fail(e, PA_E_INPUT, "piA must lie in (0, 1/2)");
    return loglik_lengths_impl(e, ids, n, lengths, piA, precision, out);
}

// One BPP_C::sample_pi step (src/PhyloAcc-GT/bpp_c.cpp:1508-1597) for a batch
// of gene-tree elements, at the full element length (no Wang-Landau split).
// Per element, on the host: the proposal -- a uniform step of half-width
// pi_delta clamped to [0.05, 0.45] with the reference's boundary terms, or,
// with resample set, a draw from Beta(prior_a + #A + #T, prior_b + #C + #G) / 2
// over the element's leaf bases; then the batched likelihood of every
// proposal in one launch; then the Metropolis decision with the Beta prior on
// 2a.  An unchanged proposal counts as accepted, as in the reference.
// Draws: stream tag PI, sweep m; the step's uniform on sub-stream 0, the
// decision's on 1, the Beta's gamma variates on 2 and 3.
extern "C" pa_status pa_sample_pi(pa_engine *e, const int32_t *ids, int32_t n, const double *lengths, int32_t m,
                                  double pi_delta, double prior_a, double prior_b, int32_t resample, int32_t precision,
                                  double *piA, double *loglik, int32_t *accepted)
{
    if (!e || !ids || !lengths || !piA || !loglik || !accepted || n < 0) return fail(e, PA_E_INPUT, "null argument");
    if (!(pi_delta > 0) || !(prior_a > 0) || !(prior_b > 0)) return fail(e, PA_E_INPUT, "pi_delta and the prior must be positive");
    pa_status st = check_ids(e, ids, n);
    if (st != PA_OK) return st;
    const Key key{e->M.key0, e->M.key1};
    const uint32_t c3 = tagword(0, TAG_PI), c2 = (uint32_t)m;
    std::vector<double> prop(n), lq(n, 0.0);     // lq: log q(cur | prop) - log q(prop | cur)
    std::vector<int32_t> run;
    for (int i = 0; i < n; i++) {
        const double a = piA[i];
        double t;
        if (resample) {
            const HostElem &E = e->elems[ids[i]];
            double cnt[4] = {0, 0, 0, 0};
            for (int s = 0; s < e->tm.S; s++)
                for (int p = 0; p < E.P; p++) {
                    const int code = E.codes[(size_t)s * E.Ppad + p];
                    if (code < 4) cnt[code] += E.w[p];
                }
            t = beta_draw(key, 2, 3, c2, c3, prior_a + cnt[0] + cnt[3], prior_b + cnt[1] + cnt[2]) / 2;
        } else {
            t = a + (2.0 * slot_uniform(key, 0, c2, c3, 0) - 1) * pi_delta;
            double lp_prop = -log(2.0 * pi_delta);
            if (t < 0.05) {
                lp_prop += log(0.05 - (a - pi_delta));
                t = 0.05;
            } else if (t > 0.45) {
                lp_prop += log((a + pi_delta) - 0.45);
                t = 0.45;
            }
            double lp_cur = -log(2.0 * pi_delta);
            if (a == 0.05 && (t - pi_delta) < 0.05) lp_cur += log(0.05 - (t - pi_delta));
            else if (a == 0.45 && (t + pi_delta) > 0.45) lp_cur += log(t + pi_delta - 0.45);
            lq[i] = lp_cur - lp_prop;
        }
        prop[i] = t;
        accepted[i] = 1;
        if (t != a) run.push_back(i);
    }
    if (run.empty()) return PA_OK;
    const int N = e->tm.N, k = (int)run.size();
    std::vector<int32_t> rid(k);
    std::vector<double> rl((size_t)k * N), rp(k), lnew(k);
    for (int j = 0; j < k; j++) {
        rid[j] = ids[run[j]];
        rp[j] = prop[run[j]];
        memcpy(&rl[(size_t)j * N], lengths + (size_t)run[j] * N, sizeof(double) * N);
    }
    if ((st = loglik_lengths_impl(e, rid.data(), k, rl.data(), rp.data(), precision, lnew.data())) != PA_OK) return st;
    auto log_beta_pdf = [](double x, double a, double b) {
        return lgamma(a + b) - lgamma(a) - lgamma(b) + (a - 1) * log(x) + (b - 1) * log1p(-x);
    };
    for (int j = 0; j < k; j++) {
        const int i = run[j];
        if (resample) {
            piA[i] = prop[i];
            loglik[i] = lnew[j];
            continue;
        }
        const double ratio = lnew[j] - loglik[i] + log_beta_pdf(2 * prop[i], prior_a, prior_b) -
                             log_beta_pdf(2 * piA[i], prior_a, prior_b) + lq[i];
        if (log(slot_uniform(key, 1, c2, c3, 0)) < ratio) {
            piA[i] = prop[i];
            loglik[i] = lnew[j];
        } else {
            accepted[i] = 0;
        }
    }
    return PA_OK;
}

// pa_loglik over device-resident inputs and outputs that allocates nothing:
// the work counter and the workgroup slots live in the library's static
// device scratch array (NOALLOC_SCRATCH_BYTES per GPU), every slot sized for
// the longest element in the arenas, and the launch is asynchronous on the
// caller's stream.  Calls that overlap on one GPU share that array and race.
extern "C" pa_status pa_loglik_noalloc(pa_engine *e, const int32_t *d_ids, int32_t n, const int32_t *d_Z,
                                       const double *d_rates, int32_t precision, double *d_out, void *stream)
{
    if (!e) return PA_E_INPUT;
    if (n < 0 || (n > 0 && (!d_ids || !d_Z || !d_rates || !d_out))) return fail(e, PA_E_INPUT, "null argument");
    if (precision != PA_FP64 && precision != PA_FP32) return fail(e, PA_E_INPUT, "precision must be PA_FP64 or PA_FP32");
    if (!e->stream || e->dev_dirty || !e->d_noalloc)
        return fail(e, PA_E_STATE, "the element arenas are not on the device: pa_loglik_noalloc uploads nothing, so "
                                   "make one device call (pa_loglik, pa_run_chains) after the last pa_elements_add");
    if (n == 0) return PA_OK;
    HIPCHK(e, hipSetDevice(e->device));
    const int S = e->tm.S, N = e->tm.N, fp64 = precision == PA_FP64;
    const size_t smem = loglik_smem_bytes(N, fp64);
    if (smem > e->lds_per_cu) return fail(e, PA_E_INPUT, "tree too large for the LDS layout");
    const int BS = (e->mean_p <= 96) ? 64 : (e->mean_p <= 320 ? 128 : 256);
    const size_t slot = loglik_slot_bytes(N, S, std::max(64, e->max_ppad), fp64);
    const size_t usable = NOALLOC_SCRATCH_BYTES - 256;
    if (slot > usable)
        return fail(e, PA_E_INPUT, "the longest element (" + std::to_string(e->max_ppad) + " padded patterns) needs " +
                                       std::to_string(slot) + " B of scratch, over the " + std::to_string(usable) +
                                       " B of the allocation-free entry point");
    const int grid = (int)std::min<size_t>({(size_t)n, (size_t)loglik_grid_max(e, fp64, BS), usable / slot});
    LoglikArgs A;
    memset(&A, 0, sizeof A);
    A.elems = e->d_elems; A.codes = e->d_codes; A.wts = e->d_w; A.tab = e->d_tab;
    A.n_elems = (int)e->elems.size();
    A.work = d_ids;
    A.n_work = n;
    A.queue = (int *)e->d_noalloc;
    A.scratch = e->d_noalloc + 256;
    A.slot_bytes = (int64_t)slot;
    A.Zin = d_Z;
    A.rates_in = d_rates;
    A.ll_out = d_out;
    hipStream_t st = stream ? (hipStream_t)stream : e->stream;
    HIPCHK(e, launch_reset(A.queue, st));          // a fill kernel, not a memset: callers may capture the stream
    HIPCHK(e, launch_loglik(e->M, A, fp64, BS, grid, st));
    return PA_OK;
}

extern "C" pa_status pa_release_scratch(pa_engine *e)
{
    if (!e) return PA_E_INPUT;
    if (!e->stream) return PA_OK;
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    release_ll_buffers(e);
    return PA_OK;
}

extern "C" pa_status pa_run_chains(pa_engine *e, const int32_t *ids, int32_t n, const pa_run_config *cfg,
                                   pa_chain_result *results, pa_traces *traces, pa_run_stats *stats)
{
    if (!e || !ids || !cfg || n < 0) return fail(e, PA_E_INPUT, "null argument");
    if (n == 0) return PA_OK;
    pa_status st = check_ids(e, ids, n);
    if (st != PA_OK) return st;
    if ((st = upload(e)) != PA_OK) return st;
    HIPCHK(e, hipSetDevice(e->device));
    const int N = e->tm.N, T = e->M.T;
    std::vector<int32_t> order = schedule(e, ids, n, cfg->mode);
    std::vector<int32_t> sorted(n);
    for (int i = 0; i < n; i++) sorted[i] = ids[order[i]];
    const bool want_tr = traces && (traces->loglik || traces->full || traces->nrate || traces->crate || traces->grate ||
                                    traces->lrate || traces->l2rate || traces->Z);
    const bool keep = cfg->keep_traces || want_tr;
    DevBuf bufs[8];
    HIPCHK(e, hipMalloc(&bufs[3].p, (size_t)n * 3 * chain::RES_D * 8));
    HIPCHK(e, hipMalloc(&bufs[4].p, (size_t)n * 3 * chain::RES_I * 4));
    HIPCHK(e, hipMalloc(&bufs[5].p, (size_t)n * 3 * N));
    if (keep) {
        HIPCHK(e, hipMalloc(&bufs[6].p, (size_t)n * 3 * 7 * T * 8));
        HIPCHK(e, hipMalloc(&bufs[7].p, (size_t)n * 3 * T * N));
    }
    chain::ChainArgs A;
    memset(&A, 0, sizeof A);
    A.res_d = (double *)bufs[3].p;
    A.res_i = (int32_t *)bufs[4].p;
    A.maxZ = (int8_t *)bufs[5].p;
    A.tr_d = keep ? (double *)bufs[6].p : nullptr;
    A.tr_Z = keep ? (int8_t *)bufs[7].p : nullptr;
    st = run_chain_kernel(e, sorted.data(), n, cfg->precision == PA_FP64, cfg->mode, cfg->block_threads, A, bufs, stats);
    if (st != PA_OK) return st;
    std::vector<double> rd((size_t)n * 3 * chain::RES_D);
    std::vector<int32_t> ri((size_t)n * 3 * chain::RES_I);
    std::vector<int8_t> mz((size_t)n * 3 * N);
    HIPCHK(e, hipMemcpy(rd.data(), bufs[3].p, rd.size() * 8, hipMemcpyDeviceToHost));
    HIPCHK(e, hipMemcpy(ri.data(), bufs[4].p, ri.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(e, hipMemcpy(mz.data(), bufs[5].p, mz.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
        const int dst = order[i];
        for (int k = 0; k < 3; k++) {
            const double *d = &rd[((size_t)i * 3 + k) * chain::RES_D];
            const int32_t *q = &ri[((size_t)i * 3 + k) * chain::RES_I];
            if (results) {
                pa_chain_result &r = results[(size_t)dst * 3 + k];
                r.max_loglik = d[0]; r.final_nrate = d[1]; r.final_crate = d[2];
                r.final_grate = d[3]; r.final_lrate = d[4]; r.final_lrate2 = d[5];
                r.max_m = q[0]; r.status = q[1]; r.accept_n = q[2]; r.accept_c = q[3]; r.flags = q[4]; r.pad = 0;
            }
            if (traces && traces->max_Z) memcpy(traces->max_Z + ((size_t)dst * 3 + k) * N, &mz[((size_t)i * 3 + k) * N], N);
        }
    }
    if (want_tr) {
        std::vector<double> td((size_t)n * 3 * 7 * T);
        HIPCHK(e, hipMemcpy(td.data(), bufs[6].p, td.size() * 8, hipMemcpyDeviceToHost));
        double *dst7[7] = {traces->loglik, traces->full, traces->nrate, traces->crate, traces->grate, traces->lrate, traces->l2rate};
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++)
                for (int f = 0; f < 7; f++)
                    if (dst7[f]) memcpy(dst7[f] + ((size_t)order[i] * 3 + k) * T, &td[(((size_t)i * 3 + k) * 7 + f) * T], (size_t)T * 8);
        if (traces->Z) {
            std::vector<int8_t> tz((size_t)n * 3 * T * N);
            HIPCHK(e, hipMemcpy(tz.data(), bufs[7].p, tz.size(), hipMemcpyDeviceToHost));
            for (int i = 0; i < n; i++)
                memcpy(traces->Z + (size_t)order[i] * 3 * T * N, &tz[(size_t)i * 3 * T * N], (size_t)3 * T * N);
        }
    }
    return PA_OK;
}

// ---------------------------------------------------- captured chain batches
//
// A prepared batched chain run: the schedule, the scratch plan and every
// device buffer are set up once, and the launch sequence -- a kernel that
// clears the work and byte counters, the persistent chain kernel, and the
// copies of the result rows into pinned host memory -- is captured once as a
// HIP graph on the engine's stream.  pa_batch_run replays it.  The chains
// draw from the same counter-based streams every time, so each replay gives
// pa_run_chains's results for the same ids and options.

struct pa_batch {
    pa_engine *e = nullptr;
    int n = 0, fp64 = 0, BS = 0, shape = 0, grid = 0, lds = 0;
    uint64_t epoch = 0;
    size_t scratch_bytes = 0;
    std::vector<int32_t> order;
    DevBuf scratch, meta, res_d, res_i, maxZ, tr_d, tr_Z;
    double *h_res_d = nullptr;
    int32_t *h_res_i = nullptr;
    unsigned long long *h_cnt = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    ~pa_batch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        (void)hipHostFree(h_res_d);
        (void)hipHostFree(h_res_i);
        (void)hipHostFree(h_cnt);
    }
};

// Enqueues the batch's launch sequence on the engine stream (under capture).
static hipError_t batch_sequence(pa_batch *b, const chain::ChainArgs &A, void *clear_at, int clear_words)
{
    pa_engine *e = b->e;
    const size_t n = (size_t)b->n;
    hipError_t err;
    if ((err = chain::launch_clear(clear_at, clear_words, e->stream)) != hipSuccess) return err;
    if ((err = chain::launch_chain(e->M, A, b->fp64, b->BS, b->grid, e->stream, b->shape)) != hipSuccess) return err;
    if ((err = hipMemcpyAsync(b->h_res_d, b->res_d.p, n * 3 * chain::RES_D * 8, hipMemcpyDeviceToHost, e->stream)) != hipSuccess)
        return err;
    if ((err = hipMemcpyAsync(b->h_res_i, b->res_i.p, n * 3 * chain::RES_I * 4, hipMemcpyDeviceToHost, e->stream)) != hipSuccess)
        return err;
    return hipMemcpyAsync(b->h_cnt, A.counters, 8, hipMemcpyDeviceToHost, e->stream);
}

extern "C" pa_status pa_batch_create(pa_engine *e, const int32_t *ids, int32_t n, const pa_run_config *cfg, pa_batch **out)
{
    if (out) *out = nullptr;
    if (!e || !ids || !cfg || !out || n <= 0) return fail(e, PA_E_INPUT, "null argument or empty batch");
    if (cfg->mode != PA_RUN_BATCHED) return fail(e, PA_E_INPUT, "a prepared batch is one batched launch (PA_RUN_BATCHED)");
    pa_status st = check_ids(e, ids, n);
    if (st != PA_OK) return st;
    if ((st = upload(e)) != PA_OK) return st;
    HIPCHK(e, hipSetDevice(e->device));
    std::unique_ptr<pa_batch> b(new pa_batch());
    b->e = e;
    b->n = n;
    b->fp64 = cfg->precision == PA_FP64;
    b->epoch = e->arena_gen;
    b->order = schedule(e, ids, n, PA_RUN_BATCHED);
    std::vector<int32_t> sorted(n);
    for (int i = 0; i < n; i++) sorted[i] = ids[b->order[i]];
    const int S = e->tm.S, N = e->tm.N;
    int BS = cfg->block_threads ? cfg->block_threads : (e->tune.block_threads ? e->tune.block_threads : pick_block(e, sorted.data(), n));
    if (BS != 64 && BS != 128 && BS != 256) return fail(e, PA_E_INPUT, "block_threads must be 64, 128 or 256");
    b->BS = BS;
    b->lds = (int)chain::chain_smem_bytes(N, S, BS, b->fp64);
    if ((size_t)b->lds > e->lds_per_cu) return fail(e, PA_E_INPUT, "tree too large for the LDS layout");
    b->shape = e->tune.shape;
    b->grid = std::min(n, chain_grid_max(e, b->fp64, BS, b->shape));
    std::vector<int> ppad(n);
    for (int i = 0; i < n; i++) ppad[i] = e->elems[sorted[i]].Ppad;
    size_t off[4];
    const size_t fixed = chain::slot_scratch_bytes(N, S, 0, b->fp64, off);
    const ScratchPlan pl = plan_scratch(ppad, b->grid, fixed, [&](int p) { return chain::chain_part_bytes(N, S, p, b->fp64); });
    chain::slot_scratch_bytes(N, S, pl.cap, b->fp64, off);
    b->scratch_bytes = pl.total;
    HIPCHK(e, hipMalloc(&b->scratch.p, pl.total));
    const size_t ids_b = ((size_t)4 * n + 15) & ~(size_t)15;
    HIPCHK(e, hipMalloc(&b->meta.p, ids_b + (size_t)8 * n + 64));
    int64_t *d_bo = (int64_t *)((char *)b->meta.p + ids_b);
    unsigned long long *d_cnt = (unsigned long long *)(d_bo + n);   // [0]: bytes, [1]: spare
    int *d_queue = (int *)(d_cnt + 2);
    HIPCHK(e, hipMemcpy(b->meta.p, sorted.data(), (size_t)4 * n, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(d_bo, pl.big_off.data(), (size_t)8 * n, hipMemcpyHostToDevice));
    HIPCHK(e, hipMalloc(&b->res_d.p, (size_t)n * 3 * chain::RES_D * 8));
    HIPCHK(e, hipMalloc(&b->res_i.p, (size_t)n * 3 * chain::RES_I * 4));
    HIPCHK(e, hipMalloc(&b->maxZ.p, (size_t)n * 3 * N));
    const int T = e->M.T;
    if (cfg->keep_traces) {        // traces stay on the device, as the reference keeps them in memory
        HIPCHK(e, hipMalloc(&b->tr_d.p, (size_t)n * 3 * 7 * T * 8));
        HIPCHK(e, hipMalloc(&b->tr_Z.p, (size_t)n * 3 * T * N));
    }
    HIPCHK(e, hipHostMalloc(&b->h_res_d, (size_t)n * 3 * chain::RES_D * 8));
    HIPCHK(e, hipHostMalloc(&b->h_res_i, (size_t)n * 3 * chain::RES_I * 4));
    HIPCHK(e, hipHostMalloc(&b->h_cnt, 16));
    HIPCHK(e, hipEventCreate(&b->t0));
    HIPCHK(e, hipEventCreate(&b->t1));
    chain::ChainArgs A;
    memset(&A, 0, sizeof A);
    A.elems = e->d_elems; A.codes = e->d_codes; A.wts = e->d_w; A.site = e->d_site; A.tab = e->d_tab;
    A.work = (const int32_t *)b->meta.p; A.n_work = n; A.queue = d_queue;
    A.scratch = (char *)b->scratch.p; A.slot_bytes = (int64_t)pl.slot;
    A.off_plin = off[0]; A.off_logt = off[1]; A.off_amb = off[2]; A.off_parts = off[3];
    A.big = (char *)b->scratch.p + pl.slots;
    A.big_off = pl.big ? d_bo : nullptr;
    A.res_d = (double *)b->res_d.p; A.res_i = (int32_t *)b->res_i.p; A.maxZ = (int8_t *)b->maxZ.p;
    A.tr_d = (double *)b->tr_d.p; A.tr_Z = (int8_t *)b->tr_Z.p;
    A.counters = d_cnt;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    HIPCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
    const hipError_t qerr = batch_sequence(b.get(), A, d_cnt, 6);   // bytes, spare, queue (+pad)
    hipGraph_t g = nullptr;
    const hipError_t cerr = hipStreamEndCapture(e->stream, &g);
    if (qerr != hipSuccess || cerr != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(e, PA_E_DEVICE, std::string("graph capture failed: ") + hipGetErrorString(qerr != hipSuccess ? qerr : cerr));
    }
    b->graph = g;
    HIPCHK(e, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    e->batches.push_back(b.get());
    *out = b.release();
    return PA_OK;
}

extern "C" pa_status pa_batch_run(pa_batch *b, pa_chain_result *results, pa_run_stats *stats)
{
    if (!b) return PA_E_INPUT;
    pa_engine *e = b->e;
    if (e->dev_dirty || e->arena_gen != b->epoch)
        return fail(e, PA_E_STATE, "elements were added after the batch was prepared; prepare it again");
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipEventRecord(b->t0, e->stream));
    HIPCHK(e, hipGraphLaunch(b->exec, e->stream));
    HIPCHK(e, hipEventRecord(b->t1, e->stream));
    HIPCHK(e, hipEventSynchronize(b->t1));
    if (results)
        for (int i = 0; i < b->n; i++)
            for (int k = 0; k < 3; k++) {
                const double *d = b->h_res_d + ((size_t)i * 3 + k) * chain::RES_D;
                const int32_t *q = b->h_res_i + ((size_t)i * 3 + k) * chain::RES_I;
                pa_chain_result &r = results[(size_t)b->order[i] * 3 + k];
                r.max_loglik = d[0]; r.final_nrate = d[1]; r.final_crate = d[2];
                r.final_grate = d[3]; r.final_lrate = d[4]; r.final_lrate2 = d[5];
                r.max_m = q[0]; r.status = q[1]; r.accept_n = q[2]; r.accept_c = q[3]; r.flags = q[4]; r.pad = 0;
            }
    if (stats) {
        float ms = 0;
        HIPCHK(e, hipEventElapsedTime(&ms, b->t0, b->t1));
        stats->elapsed_ms = ms;
        stats->algorithmic_bytes = (double)*b->h_cnt;
        stats->grid = b->grid;
        stats->block_threads = b->BS;
        stats->lds_bytes = b->lds;
        stats->n_launches = 1;
        stats->scratch_bytes = (double)b->scratch_bytes;
    }
    return PA_OK;
}

static void batch_free(pa_batch *b)
{
    pa_engine *e = b->e;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    e->batches.erase(std::remove(e->batches.begin(), e->batches.end(), b), e->batches.end());
    delete b;
}

// Batches still alive when their engine is destroyed are freed with it.
extern "C" void pa_batch_destroy(pa_batch *b)
{
    if (b) batch_free(b);
}

// ------------------------------------------------------- chain summaries

namespace {

double lnbeta(double a, double b) { return lgamma(a) + lgamma(b) - lgamma(a + b); }

double log_gamma_density(double x, double a, double b)
{
    if (x < 0) return -INFINITY;
    if (x == 0) return a == 1 ? -log(b) : (a > 1 ? -INFINITY : INFINITY);
    return (a - 1) * log(x) - x / b - lgamma(a) - a * log(b);
}

}  // namespace

// Eval2 (src/PhyloAcc-ST/bpp_c2.cpp:14-353) and Output_init
// (bpp_c2.cpp:739-785) of one element's model chain from its traces.
extern "C" pa_status pa_chain_summary(pa_engine *e, int32_t id, int32_t model, const double *full,
                                      const double *nrate, const double *crate, const double *grate,
                                      const double *lrate, const double *l2rate, const int8_t *Z,
                                      const int8_t *max_Z, int32_t precision, double *vb, double *medians,
                                      double *postZ)
{
    if (!e || id < 0 || id >= (int)e->elems.size() || model < 0 || model > 2) return fail(e, PA_E_INPUT, "bad argument");
    const HostElem &E = e->elems[id];
    if (E.filtered) return fail(e, PA_E_INPUT, "element is filtered");
    const pa_options &o = e->opt;
    const int N = e->tm.N, T = o.burnin + o.mcmc, nb = o.burnin, nm = o.mcmc;
    const int resZ = model == 0 ? 0 : (model == 1 ? 2 : 1);
    const int32_t *tab = E.tab.data();
    const std::vector<int> &nodes = E.nodes;
    const int root = E.root;
    auto par = [&](int s) { return tab[tab_par(N) + s]; };
    auto flag = [&](int s) { return (uint32_t)tab[tab_flags(N) + s]; };
    std::vector<int> fixZ(N, 0);
    for (int s : nodes) {
        int fz = (flag(s) & NF_UPPER) || s == root;
        if (resZ == 2 && (flag(s) & NF_UPPER_CONS)) fz = 1;
        if (resZ == 0 && s != root) fz = 1;
        fixZ[s] = fz;
    }
    if (medians) {
        const double *tr[5] = {nrate, crate, grate, lrate, l2rate};
        std::vector<double> tmp(T);
        for (int k = 0; k < 5; k++) {
            if (!tr[k]) { medians[k] = NAN; continue; }
            tmp.assign(tr[k], tr[k] + T);
            std::sort(tmp.begin(), tmp.end());
            medians[k] = tmp[T / 2];
        }
    }
    if (postZ) {
        for (int s = 0; s < N; s++) {
            int cnt[4] = {0, 0, 0, 0};
            for (int i = nb; i < T; i++) cnt[Z[(size_t)i * N + s] + 1]++;
            if (flag(s) & NF_MISSING) cnt[0] = nm;   // missing nodes: column 0 is MCMC (bpp_c2.cpp:768-771)
            for (int k = 0; k < 4; k++) postZ[s * 4 + k] = (double)cnt[k] / nm;
        }
    }
    if (!vb) return PA_OK;
    // configurations of the uncertain nodes after burn-in, newest first, MaxZ first
    std::vector<int> unc;
    for (int s = 0; s < N; s++) {
        int agree = 0;
        for (int i = nb; i < T; i++) agree += Z[(size_t)i * N + s] == max_Z[s];
        if (agree < nm) unc.push_back(s);
    }
    std::vector<const int8_t *> conf{max_Z};
    std::vector<int> num{0};
    std::vector<double> vc{0}, vn{0}, mc{0}, mn{0}, cov{0}, sfl{0};
    for (int i = T - 1; i >= nb; i--) {
        const int8_t *row = Z + (size_t)i * N;
        size_t j = 0;
        for (; j < conf.size(); j++) {
            bool same = true;
            for (int s : unc) if (row[s] != conf[j][s]) { same = false; break; }
            if (same) break;
        }
        if (j == conf.size()) {
            conf.push_back(row);
            num.push_back(0); vc.push_back(0); vn.push_back(0); mc.push_back(0); mn.push_back(0);
            cov.push_back(0); sfl.push_back(0);
        }
        num[j]++;
        vc[j] += crate[i] * crate[i];
        mc[j] += crate[i];
        if (resZ != 0) { vn[j] += nrate[i] * nrate[i]; mn[j] += nrate[i]; cov[j] += crate[i] * nrate[i]; }
        sfl[j] += full[i] / nm;
    }
    const double lp[3] = {log(0.5), log(0.5), -INFINITY};
    auto priorZ = [&](const int8_t *cz) {
        double n[3][3] = {{0}};
        for (size_t k = 0; k + 1 < nodes.size(); k++) n[cz[nodes[k]]][cz[par(nodes[k])]] += 1;
        double p = lnbeta(o.hyper_grate_a + n[1][0] + n[2][0], o.hyper_grate_b + n[0][0]) - lnbeta(o.hyper_grate_a, o.hyper_grate_b);
        for (size_t k = 0; k + 1 < nodes.size(); k++)
            if (fixZ[nodes[k]]) n[cz[nodes[k]]][cz[par(nodes[k])]] -= 1;
        p += lnbeta(o.hyper_lrate_a + n[2][1], o.hyper_lrate_b + n[1][1]) - lnbeta(o.hyper_lrate_a, o.hyper_lrate_b);
        if (o.hyper_lrate2_a > 0)
            p += lnbeta(o.hyper_lrate2_a + n[1][2], o.hyper_lrate2_b + n[2][2]) - lnbeta(o.hyper_lrate2_a, o.hyper_lrate2_b);
        return p + lp[cz[root]];
    };
    double VB = 0;
    int eff = 0;
    for (size_t i = 0; i < conf.size(); i++) {
        const double w = (double)num[i] / nm;
        if (w < 0.02) continue;
        sfl[i] += num[i] * priorZ(conf[i]) / nm;
        double logdet;
        if (resZ == 0) logdet = log(vc[i] - mc[i] * mc[i] / num[i]) - log((double)num[i]);
        else {
            const double a = vc[i] - mc[i] * mc[i] / num[i], b = vn[i] - mn[i] * mn[i] / num[i];
            const double cv = cov[i] - mn[i] * mc[i] / num[i];
            logdet = log(a * b - cv * cv) - 2 * log((double)num[i]);
        }
        if (std::isnan(logdet)) continue;
        eff += num[i];
        VB += -w * log((double)num[i]) + sfl[i] + 0.5 * logdet * num[i] / nm;
    }
    const double TWO_PI = 6.283185307179586;
    if (eff == 0) {
        // rates and Z treated as independent: P(X | Z_i, r*) for every configuration at r* = posterior means
        double varn = 0, varc = 0, cv = 0, meann = 0, meanc = 0;
        for (size_t i = 0; i < conf.size(); i++) {
            varc += vc[i]; meanc += mc[i];
            if (resZ != 0) { meann += mn[i]; varn += vn[i]; cv += cov[i]; }
        }
        const double rn = meann / nm, rc = meanc / nm;
        const int nc = (int)conf.size();
        std::vector<int32_t> ids(nc, id), Zs((size_t)nc * N);
        std::vector<double> rates((size_t)nc * 2), ll(nc);
        for (int i = 0; i < nc; i++) {
            for (int s = 0; s < N; s++) Zs[(size_t)i * N + s] = conf[i][s];
            rates[2 * i] = rc;
            rates[2 * i + 1] = resZ != 0 ? rn : 1.0;
        }
        pa_status st = pa_loglik(e, ids.data(), nc, Zs.data(), rates.data(), precision, ll.data());
        if (st != PA_OK) return st;
        double mx = -INFINITY;
        for (int i = 0; i < nc; i++) { sfl[i] = ll[i] + priorZ(conf[i]); mx = std::max(mx, sfl[i]); }
        double acc = 0;
        for (int i = 0; i < nc; i++) acc += exp(sfl[i] - mx);
        VB = mx == -INFINITY ? -INFINITY : log(acc) + mx;
        VB += log_gamma_density(rc, o.cprior_a, o.cprior_b);
        if (resZ != 0) VB += log_gamma_density(rn, o.nprior_a, o.nprior_b);
        double logdet;
        if (resZ == 0) {
            varc -= meanc * meanc / nm;
            logdet = log(varc) - log((double)nm);
            VB += 0.5 * logdet + 0.5 * log(TWO_PI);
        } else {
            varn -= meann * meann / nm; varc -= meanc * meanc / nm; cv -= meann * meanc / nm;
            logdet = log(varc * varn - cv * cv) - 2 * log((double)nm);
            VB += 0.5 * logdet + log(TWO_PI);
        }
    } else if (resZ != 0) {
        VB = VB * ((double)nm / eff) + log((double)nm) + log(TWO_PI) + 1;
    } else {
        VB = VB * ((double)nm / eff) + log((double)nm) + 0.5 * (log(TWO_PI) + 1);
    }
    *vb = VB;
    return PA_OK;
}

// ------------------------------------------------------------- autotuner
//
// Batched chain launches use one block size and one pattern-loop shape.
// pa_autotune times each pair on the batch itself with short chains and
// keeps the fastest for later calls with block_threads = 0.

static void shape_dims(int fp64, int shape, int32_t *U, int32_t *K)
{
    static const int u32[3] = {2, 1, 2}, k32[3] = {4, 4, 8}, u64[3] = {1, 2, 1}, k64[3] = {2, 2, 4};
    *U = fp64 ? u64[shape] : u32[shape];
    *K = fp64 ? k64[shape] : k32[shape];
}

extern "C" pa_status pa_set_tuning(pa_engine *e, const pa_tuning *t)
{
    if (!e || !t) return fail(e, PA_E_INPUT, "null argument");
    if (t->block_threads != 0 && t->block_threads != 64 && t->block_threads != 128 && t->block_threads != 256)
        return fail(e, PA_E_INPUT, "tuning: block_threads must be 0, 64, 128 or 256");
    if (t->shape < 0 || t->shape >= chain::N_SHAPES) return fail(e, PA_E_INPUT, "tuning: shape out of range");
    e->tune = *t;
    shape_dims(0, t->shape, &e->tune.round_nodes, &e->tune.chunk_nodes);
    return PA_OK;
}

extern "C" pa_status pa_get_tuning(pa_engine *e, pa_tuning *t)
{
    if (!e || !t) return fail(e, PA_E_INPUT, "null argument");
    *t = e->tune;
    return PA_OK;
}

extern "C" pa_status pa_autotune(pa_engine *e, const int32_t *ids, int32_t n, int32_t precision, int32_t sample,
                                 int32_t states, pa_tuning *out)
{
    if (!e || !ids || n <= 0) return fail(e, PA_E_INPUT, "null argument");
    pa_status st = check_ids(e, ids, n);
    if (st != PA_OK) return st;
    if ((st = upload(e)) != PA_OK) return st;
    if (states <= 1) states = 8;
    // the batch itself (or an evenly spaced subset of it, by pattern count)
    std::vector<int32_t> smp(ids, ids + n);
    if (sample > 0 && sample < n) {
        std::stable_sort(smp.begin(), smp.end(), [&](int a, int b) { return e->elems[a].P < e->elems[b].P; });
        std::vector<int32_t> sub;
        for (int i = 0; i < sample; i++) sub.push_back(smp[(size_t)((i + 0.5) * n / sample)]);
        smp.swap(sub);
    }
    const std::vector<int32_t> order = schedule(e, smp.data(), (int32_t)smp.size(), PA_RUN_BATCHED);
    std::vector<int32_t> sorted(smp.size());
    for (size_t i = 0; i < smp.size(); i++) sorted[i] = smp[order[i]];
    const int N = e->tm.N, m = (int)sorted.size();
    const ModelDev saved = e->M;
    e->M.T = states;
    e->M.burnin = states / 4;
    e->M.mcmc = states - states / 4;
    pa_tuning best = {0, 0, 0, 0, -1.0};
    const int cand[3] = {64, 128, 256};
    const pa_tuning keep = e->tune;
    // one timed launch of `T`-state chains at block size bs and shape sh
    auto time_one = [&](int bs, int sh, int T, double *rate) -> pa_status {
        DevBuf bufs[8];
        pa_run_stats rs;
        chain::ChainArgs A;
        memset(&A, 0, sizeof A);
        if (hipMalloc(&bufs[3].p, (size_t)m * 3 * chain::RES_D * 8) != hipSuccess ||
            hipMalloc(&bufs[4].p, (size_t)m * 3 * chain::RES_I * 4) != hipSuccess ||
            hipMalloc(&bufs[5].p, (size_t)m * 3 * N) != hipSuccess) return fail(e, PA_E_OOM, "autotune buffers");
        A.res_d = (double *)bufs[3].p;
        A.res_i = (int32_t *)bufs[4].p;
        A.maxZ = (int8_t *)bufs[5].p;
        e->tune.shape = sh;
        e->M.T = T;
        e->M.burnin = T / 4;
        e->M.mcmc = T - T / 4;
        pa_status r = run_chain_kernel(e, sorted.data(), m, precision == PA_FP64, PA_RUN_BATCHED, bs, A, bufs, &rs);
        if (r == PA_OK) *rate = (double)m * 3 * (T - 1) / (rs.elapsed_ms / 1e3);
        return r;
    };
    // every (block, shape) once, then the three fastest again at twice the
    // chain length, twice each, keeping each one's better time: neighbours
    // that differ by a percent or two are then told apart reliably
    std::vector<std::pair<double, std::pair<int, int>>> first;
    for (int a = 0; a < 3 && st == PA_OK; a++)
        for (int sh = 0; sh < chain::N_SHAPES && st == PA_OK; sh++) {
            double rate = 0;
            if ((st = time_one(cand[a], sh, states, &rate)) == PA_OK) first.push_back({rate, {cand[a], sh}});
        }
    std::sort(first.begin(), first.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
    for (size_t q = 0; q < first.size() && q < 3 && st == PA_OK; q++) {
        double r = 0;
        for (int rep = 0; rep < 2 && st == PA_OK; rep++) {
            double x = 0;
            if ((st = time_one(first[q].second.first, first[q].second.second, 2 * states, &x)) == PA_OK) r = std::max(r, x);
        }
        if (st == PA_OK && r > best.rate) {
            best = {first[q].second.first, first[q].second.second, 0, 0, r};
            shape_dims(precision == PA_FP64, best.shape, &best.round_nodes, &best.chunk_nodes);
        }
    }
    e->tune = keep;
    e->M = saved;
    if (st != PA_OK) return st;
    e->tune = best;
    if (out) *out = best;
    return PA_OK;
}

// ------------------------------------------------------------ serving loop
//
// Continuous request loop.  Requests wait in arrival order.  Up to `lanes`
// launches are in flight at once, each on a stream of its own with its own
// work counter, and each holds a range of the server's HBM pool (first fit)
// for its scratch while it runs.  A free lane takes the waiting requests that
// fit the largest free range (up to max_chains) and launches the chain kernel
// over them, then the log-likelihood kernel over the waiting own-tree
// requests.  Elements retire one by one: the kernel writes an element's
// result rows into host-visible memory and then its completion flag, and
// every step collects the flags that are set, so a short chain does not wait
// for the longest one of its launch.  A launch ends with its last element;
// its pool range is then free for the next admission while the other lanes
// keep running, so requests are admitted while earlier launches still occupy
// the device and the hardware gives their workgroups the slots that the
// retiring launches free.  Each element's chain is run by one workgroup from
// its own random streams and the block size is fixed per server, so a result
// does not depend on the lane or on what else is in flight.

namespace {

struct Request {
    int64_t ticket;
    int32_t eid;
    double arrival, admit;
};

struct LRequest {
    int64_t ticket;
    int32_t eid;
    double arrival, admit;
    std::vector<double> len;
};

// First-fit allocator over [0, cap) in 256-byte units; free ranges coalesce.
struct PoolMap {
    size_t cap = 0, used = 0;
    std::map<size_t, size_t> free_;   // offset -> length
    static size_t round(size_t n) { return (n + 255) & ~(size_t)255; }
    void init(size_t c)
    {
        cap = c & ~(size_t)255;
        used = 0;
        free_.clear();
        if (cap) free_[0] = cap;
    }
    size_t largest() const
    {
        size_t m = 0;
        for (const auto &f : free_) m = std::max(m, f.second);
        return m;
    }
    bool alloc(size_t n, size_t *off)
    {
        n = round(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const size_t o = it->first, rest = it->second - n;
            free_.erase(it);
            if (rest) free_[o + n] = rest;
            *off = o;
            used += n;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = round(n);
        if (!n) return;
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

// One launch slot of the server: a stream, a work counter, work lists and
// result rows for up to max_chains requests of each kind.
struct Lane {
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int *d_queue = nullptr;
    bool busy = false;
    size_t pool_off = 0, pool_len = 0;
    int32_t *d_ids = nullptr, *h_ids = nullptr;     // chain work list
    int64_t *d_bo = nullptr, *h_bo = nullptr;       // long elements' own regions
    double *h_res_d = nullptr, *v_res_d = nullptr;  // [max][3][RES_D], host-visible (h_) and its device view (v_)
    int32_t *h_res_i = nullptr, *v_res_i = nullptr; // [max][3][RES_I]
    int32_t *h_done = nullptr, *v_done = nullptr;   // [max] completion flags
    int8_t *d_maxZ = nullptr;                       // [max][3][N]
    int32_t *d_lids = nullptr, *h_lids = nullptr;
    int64_t *d_lbo = nullptr, *h_lbo = nullptr;
    double *d_llen = nullptr, *h_llen = nullptr, *d_lout = nullptr, *h_lout = nullptr;
    std::vector<Request> batch;                     // launch order
    std::vector<LRequest> lbatch;
    std::vector<uint8_t> retired;
    int n_retired = 0;
};

}  // namespace

struct pa_server {
    pa_engine *e = nullptr;
    int fp64 = 0, BS = 128, shape = 0, max_chains = 4096;
    size_t pool_bytes = 0;
    char *d_pool = nullptr;
    PoolMap pool;
    std::vector<Lane> lanes;
    std::deque<Request> waiting;
    std::deque<LRequest> lwait;
    std::deque<pa_serve_result> finished;
    int64_t next_ticket = 0;
    int in_flight = 0;                              // admitted, not yet retired
    pa_serve_stats st{};
};

extern "C" double pa_serve_clock(void)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void lane_free(Lane &L)
{
    if (L.st) (void)hipStreamSynchronize(L.st);
    (void)hipFree(L.d_queue); (void)hipFree(L.d_ids); (void)hipFree(L.d_bo); (void)hipFree(L.d_maxZ);
    (void)hipFree(L.d_lids); (void)hipFree(L.d_lbo); (void)hipFree(L.d_llen); (void)hipFree(L.d_lout);
    (void)hipHostFree(L.h_ids); (void)hipHostFree(L.h_bo); (void)hipHostFree(L.h_res_d); (void)hipHostFree(L.h_res_i);
    (void)hipHostFree(L.h_done); (void)hipHostFree(L.h_lids); (void)hipHostFree(L.h_lbo); (void)hipHostFree(L.h_llen);
    (void)hipHostFree(L.h_lout);
    if (L.ev0) (void)hipEventDestroy(L.ev0);
    if (L.ev1) (void)hipEventDestroy(L.ev1);
    if (L.st) (void)hipStreamDestroy(L.st);
    L = Lane();
}

static void serve_free(pa_server *s)
{
    for (Lane &L : s->lanes) lane_free(L);
    (void)hipFree(s->d_pool);
    s->d_pool = nullptr;
}

static pa_status lane_alloc(pa_engine *e, Lane &L, size_t M, int N)
{
    const unsigned vis = hipHostMallocMapped | hipHostMallocCoherent;
    HIPCHK(e, hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking));
    HIPCHK(e, hipEventCreate(&L.ev0));
    HIPCHK(e, hipEventCreate(&L.ev1));
    HIPCHK(e, hipMalloc(&L.d_queue, 256));
    HIPCHK(e, hipMalloc(&L.d_ids, sizeof(int32_t) * M));
    HIPCHK(e, hipMalloc(&L.d_bo, sizeof(int64_t) * M));
    HIPCHK(e, hipMalloc(&L.d_maxZ, (size_t)3 * N * M));
    HIPCHK(e, hipMalloc(&L.d_lids, sizeof(int32_t) * M));
    HIPCHK(e, hipMalloc(&L.d_lbo, sizeof(int64_t) * M));
    HIPCHK(e, hipMalloc(&L.d_llen, sizeof(double) * M * N));
    HIPCHK(e, hipMalloc(&L.d_lout, sizeof(double) * M));
    HIPCHK(e, hipHostMalloc(&L.h_ids, sizeof(int32_t) * M));
    HIPCHK(e, hipHostMalloc(&L.h_bo, sizeof(int64_t) * M));
    HIPCHK(e, hipHostMalloc(&L.h_res_d, sizeof(double) * 3 * chain::RES_D * M, vis));
    HIPCHK(e, hipHostMalloc(&L.h_res_i, sizeof(int32_t) * 3 * chain::RES_I * M, vis));
    HIPCHK(e, hipHostMalloc(&L.h_done, sizeof(int32_t) * M, vis));
    HIPCHK(e, hipHostGetDevicePointer((void **)&L.v_res_d, L.h_res_d, 0));
    HIPCHK(e, hipHostGetDevicePointer((void **)&L.v_res_i, L.h_res_i, 0));
    HIPCHK(e, hipHostGetDevicePointer((void **)&L.v_done, L.h_done, 0));
    HIPCHK(e, hipHostMalloc(&L.h_lids, sizeof(int32_t) * M));
    HIPCHK(e, hipHostMalloc(&L.h_lbo, sizeof(int64_t) * M));
    HIPCHK(e, hipHostMalloc(&L.h_llen, sizeof(double) * M * N));
    HIPCHK(e, hipHostMalloc(&L.h_lout, sizeof(double) * M));
    return PA_OK;
}

extern "C" pa_status pa_serve_create(pa_engine *e, const pa_serve_config *cfg, pa_server **out)
{
    if (out) *out = nullptr;
    if (!e || !cfg || !out) return fail(e, PA_E_INPUT, "null argument");
    if (cfg->lanes < 0 || cfg->lanes > PA_SERVE_MAX_LANES) return fail(e, PA_E_INPUT, "lanes must be 0..PA_SERVE_MAX_LANES");
    pa_status st = ensure_device(e);
    if (st != PA_OK) return st;
    HIPCHK(e, hipSetDevice(e->device));
    std::unique_ptr<pa_server> s(new pa_server());
    s->e = e;
    s->fp64 = cfg->precision == PA_FP64;
    if ((st = upload(e)) != PA_OK) return st;
    s->BS = cfg->block_threads ? cfg->block_threads
                               : (e->tune.block_threads ? e->tune.block_threads : block_for_mean(e->mean_p));
    s->shape = e->tune.shape;
    if (s->BS != 64 && s->BS != 128 && s->BS != 256) return fail(e, PA_E_INPUT, "block_threads must be 64, 128 or 256");
    const int N = e->tm.N, S = e->tm.S;
    if (chain::chain_smem_bytes(N, S, s->BS, s->fp64) > e->lds_per_cu || loglik_smem_bytes(N, s->fp64) > e->lds_per_cu)
        return fail(e, PA_E_INPUT, "tree too large for the LDS layout");
    s->max_chains = cfg->max_chains > 0 ? cfg->max_chains : 4096;
    s->pool_bytes = cfg->hbm_budget_bytes > 0 ? (size_t)cfg->hbm_budget_bytes : ((size_t)4 << 30);
    const size_t M = (size_t)s->max_chains;
    struct Alloc { pa_server *s; ~Alloc() { if (s) serve_free(s); } } guard{s.get()};
    HIPCHK(e, hipMalloc(&s->d_pool, s->pool_bytes));
    s->pool.init(s->pool_bytes);
    s->lanes.resize(cfg->lanes ? cfg->lanes : PA_SERVE_DEFAULT_LANES);
    for (Lane &L : s->lanes)
        if ((st = lane_alloc(e, L, M, N)) != PA_OK) return st;
    guard.s = nullptr;
    *out = s.release();
    return PA_OK;
}

extern "C" pa_status pa_serve_submit(pa_server *s, const int32_t *ids, int32_t n, const double *arrival_s, int64_t *tickets)
{
    if (!s || (!ids && n > 0) || n < 0) return PA_E_INPUT;
    pa_engine *e = s->e;
    pa_status st = check_ids(e, ids, n);
    if (st != PA_OK) return st;
    const double now = pa_serve_clock();
    for (int i = 0; i < n; i++) {
        const int64_t t = s->next_ticket++;
        s->waiting.push_back({t, ids[i], arrival_s ? arrival_s[i] : now, 0.0});
        if (tickets) tickets[i] = t;
    }
    s->st.submitted += n;
    return PA_OK;
}

extern "C" pa_status pa_serve_submit_lengths(pa_server *s, const int32_t *ids, int32_t n, const double *lengths,
                                             const double *arrival_s, int64_t *tickets)
{
    if (!s || (!ids && n > 0) || (!lengths && n > 0) || n < 0) return PA_E_INPUT;
    pa_engine *e = s->e;
    pa_status st = check_ids(e, ids, n);
    if (st != PA_OK) return st;
    const int N = e->tm.N;
    const double now = pa_serve_clock();
    for (int i = 0; i < n; i++) {
        const int64_t t = s->next_ticket++;
        s->lwait.push_back({t, ids[i], arrival_s ? arrival_s[i] : now, 0.0,
                            std::vector<double>(lengths + (size_t)i * N, lengths + (size_t)(i + 1) * N)});
        if (tickets) tickets[i] = t;
    }
    s->st.submitted += n;
    return PA_OK;
}

// A server's slots hold elements of up to SERVE_SLOT_PATTERNS padded
// patterns; longer ones get a partial region of their own in the launch's
// scratch, so one long request does not size every slot.
static const int SERVE_SLOT_PATTERNS = 512;

static size_t chain_slot(pa_server *s)
{
    size_t off[4];
    return chain::slot_scratch_bytes(s->e->tm.N, s->e->tm.S, SERVE_SLOT_PATTERNS, s->fp64, off);
}
static size_t chain_big(pa_server *s, int ppad)
{
    return ppad > SERVE_SLOT_PATTERNS ? chain::chain_part_bytes(s->e->tm.N, s->e->tm.S, ppad, s->fp64) : 0;
}
static size_t loglik_slot(pa_server *s) { return loglik_slot_bytes(s->e->tm.N, s->e->tm.S, SERVE_SLOT_PATTERNS, s->fp64); }
static size_t loglik_big(pa_server *s, int ppad)
{
    return ppad > SERVE_SLOT_PATTERNS ? loglik_slot_bytes(s->e->tm.N, s->e->tm.S, ppad, s->fp64) : 0;
}

static void finish_unfit(pa_server *s, int64_t ticket, int32_t eid, int kind, double arrival, double now)
{
    pa_serve_result out;
    memset(&out, 0, sizeof out);
    out.ticket = ticket; out.element = eid; out.status = PA_ELEM_NUMERIC_FAIL; out.kind = kind;
    out.arrival_s = arrival; out.admit_s = now; out.finish_s = now;
    s->finished.push_back(out);
    s->st.finished++;
}

// Admits waiting requests (arrival order) into free lane L within the pool's
// largest free range and launches them.  Requests whose scratch alone exceeds
// the whole pool finish at once with status PA_ELEM_NUMERIC_FAIL.
static pa_status lane_launch(pa_server *s, Lane &L, bool others_busy)
{
    pa_engine *e = s->e;
    const int N = e->tm.N, S = e->tm.S;
    const int gmax_c = chain_grid_max(e, s->fp64, s->BS, s->shape);
    const int gmax_l = loglik_grid_max(e, s->fp64, s->BS);
    const size_t room = s->pool.largest();
    const double now = pa_serve_clock();
    // chain requests: slots for min(k, grid) workgroups plus the long elements' own regions
    size_t bigC = 0;
    L.batch.clear();
    std::vector<int64_t> boC;
    while (!s->waiting.empty() && (int)L.batch.size() < s->max_chains) {
        const Request &r = s->waiting.front();
        const size_t b = chain_big(s, e->elems[r.eid].Ppad);
        const int k = (int)L.batch.size() + 1;
        const size_t need = PoolMap::round(chain_slot(s) * (size_t)std::min(k, gmax_c) + bigC + b);
        if (need > room) {
            if (!L.batch.empty() || others_busy || need <= s->pool.cap) break;   // waits for room
            finish_unfit(s, r.ticket, r.eid, PA_SERVE_CHAINS, r.arrival, now);    // cannot run even alone
            s->waiting.pop_front();
            continue;
        }
        boC.push_back(b ? (int64_t)bigC : -1);
        bigC += b;
        L.batch.push_back(r);
        L.batch.back().admit = now;
        s->waiting.pop_front();
    }
    const int nc = (int)L.batch.size();
    const int gc = std::min(nc, gmax_c);
    const size_t slotsC = nc ? chain_slot(s) * (size_t)gc : 0;
    const size_t scr_c = nc ? PoolMap::round(slotsC + bigC) : 0;
    // log-likelihood requests, after the chains' scratch
    size_t bigL = 0;
    L.lbatch.clear();
    std::vector<int64_t> boL;
    while (!s->lwait.empty() && (int)L.lbatch.size() < s->max_chains) {
        const LRequest &r = s->lwait.front();
        const size_t b = loglik_big(s, e->elems[r.eid].Ppad);
        const int k = (int)L.lbatch.size() + 1;
        const size_t alone = PoolMap::round(loglik_slot(s) + b);
        if (scr_c + PoolMap::round(loglik_slot(s) * (size_t)std::min(k, gmax_l) + bigL + b) > room) {
            if (!L.lbatch.empty() || nc || others_busy || alone <= s->pool.cap) break;
            finish_unfit(s, r.ticket, r.eid, PA_SERVE_LOGLIK, r.arrival, now);
            s->lwait.pop_front();
            continue;
        }
        boL.push_back(b ? (int64_t)bigL : -1);
        bigL += b;
        L.lbatch.push_back(std::move(s->lwait.front()));
        L.lbatch.back().admit = now;
        s->lwait.pop_front();
    }
    const int nl = (int)L.lbatch.size();
    const int gl = std::min(nl, gmax_l);
    const size_t slotsL = nl ? loglik_slot(s) * (size_t)gl : 0;
    if (nc == 0 && nl == 0) return PA_OK;
    L.pool_len = scr_c + (nl ? PoolMap::round(slotsL + bigL) : 0);
    if (!s->pool.alloc(L.pool_len, &L.pool_off))   // cannot happen: the batch was sized to the largest range
        return fail(e, PA_E_STATE, "serving pool allocation failed");
    char *base = s->d_pool + L.pool_off;
    L.retired.assign(nc, 0);
    L.n_retired = 0;
    HIPCHK(e, hipEventRecord(L.ev0, L.st));
    if (nc) {
        for (int i = 0; i < nc; i++) L.h_ids[i] = L.batch[i].eid;
        memcpy(L.h_bo, boC.data(), sizeof(int64_t) * nc);
        memset(L.h_done, 0, sizeof(int32_t) * nc);   // the host owns the flags until the launch below
        HIPCHK(e, hipMemcpyAsync(L.d_ids, L.h_ids, sizeof(int32_t) * nc, hipMemcpyHostToDevice, L.st));
        HIPCHK(e, hipMemcpyAsync(L.d_bo, L.h_bo, sizeof(int64_t) * nc, hipMemcpyHostToDevice, L.st));
        chain::ChainArgs A;
        memset(&A, 0, sizeof A);
        size_t off[4];
        const size_t slot = chain::slot_scratch_bytes(N, S, SERVE_SLOT_PATTERNS, s->fp64, off);
        A.elems = e->d_elems; A.codes = e->d_codes; A.wts = e->d_w; A.site = e->d_site; A.tab = e->d_tab;
        A.work = L.d_ids; A.n_work = nc; A.queue = L.d_queue;
        A.scratch = base; A.slot_bytes = (int64_t)slot;
        A.off_plin = off[0]; A.off_logt = off[1]; A.off_amb = off[2]; A.off_parts = off[3];
        A.big = base + slotsC;
        A.big_off = bigC ? L.d_bo : nullptr;
        A.res_d = L.v_res_d; A.res_i = L.v_res_i; A.maxZ = L.d_maxZ;
        A.done = L.v_done;
        HIPCHK(e, chain::launch_reset(L.d_queue, L.st));
        HIPCHK(e, chain::launch_chain(e->M, A, s->fp64, s->BS, gc, L.st, s->shape));
    }
    if (nl) {
        for (int i = 0; i < nl; i++) {
            L.h_lids[i] = L.lbatch[i].eid;
            memcpy(L.h_llen + (size_t)i * N, L.lbatch[i].len.data(), sizeof(double) * N);
        }
        memcpy(L.h_lbo, boL.data(), sizeof(int64_t) * nl);
        HIPCHK(e, hipMemcpyAsync(L.d_lids, L.h_lids, sizeof(int32_t) * nl, hipMemcpyHostToDevice, L.st));
        HIPCHK(e, hipMemcpyAsync(L.d_lbo, L.h_lbo, sizeof(int64_t) * nl, hipMemcpyHostToDevice, L.st));
        HIPCHK(e, hipMemcpyAsync(L.d_llen, L.h_llen, sizeof(double) * (size_t)nl * N, hipMemcpyHostToDevice, L.st));
        LoglikArgs A;
        memset(&A, 0, sizeof A);
        A.elems = e->d_elems; A.codes = e->d_codes; A.wts = e->d_w; A.tab = e->d_tab;
        A.n_elems = (int)e->elems.size();
        A.work = L.d_lids; A.n_work = nl; A.queue = L.d_queue;
        A.scratch = base + scr_c; A.slot_bytes = (int64_t)loglik_slot(s);
        A.big = base + scr_c + slotsL;
        A.big_off = bigL ? L.d_lbo : nullptr;
        A.ll_out = L.d_lout; A.len_in = L.d_llen;
        HIPCHK(e, launch_reset(L.d_queue, L.st));
        HIPCHK(e, launch_loglik(e->M, A, s->fp64, s->BS, gl, L.st));
        HIPCHK(e, hipMemcpyAsync(L.h_lout, L.d_lout, sizeof(double) * nl, hipMemcpyDeviceToHost, L.st));
    }
    HIPCHK(e, hipEventRecord(L.ev1, L.st));
    L.busy = true;
    s->in_flight += nc + nl;
    s->st.admitted += nc + nl;
    s->st.launches++;
    return PA_OK;
}

// Collects the elements of busy lanes whose completion flag is set; a lane
// whose launch has ended retires its log-likelihood requests and returns its
// pool range.  block: wait for every busy lane's launch to end.
static pa_status serve_collect(pa_server *s, bool block, int32_t *n_finished)
{
    pa_engine *e = s->e;
    for (Lane &L : s->lanes) {
        if (!L.busy) continue;
        bool ended = false;
        if (block) {
            HIPCHK(e, hipEventSynchronize(L.ev1));
            ended = true;
        } else {
            const hipError_t q = hipEventQuery(L.ev1);
            if (q != hipErrorNotReady) { HIPCHK(e, q); ended = true; }
        }
        const double now = pa_serve_clock();
        const int nc = (int)L.batch.size();
        for (int i = 0; i < nc; i++) {
            if (L.retired[i] || !__atomic_load_n(L.h_done + i, __ATOMIC_ACQUIRE)) continue;
            const Request &r = L.batch[i];
            pa_serve_result out;
            memset(&out, 0, sizeof out);
            out.ticket = r.ticket; out.element = r.eid; out.kind = PA_SERVE_CHAINS;
            out.arrival_s = r.arrival; out.admit_s = r.admit; out.finish_s = now;
            int bad = 0;
            for (int k = 0; k < 3; k++) {
                const double *d = L.h_res_d + ((size_t)i * 3 + k) * chain::RES_D;
                const int32_t *q = L.h_res_i + ((size_t)i * 3 + k) * chain::RES_I;
                pa_chain_result &c = out.result[k];
                c.max_loglik = d[0]; c.final_nrate = d[1]; c.final_crate = d[2];
                c.final_grate = d[3]; c.final_lrate = d[4]; c.final_lrate2 = d[5];
                c.max_m = q[0]; c.status = q[1]; c.accept_n = q[2]; c.accept_c = q[3]; c.flags = q[4]; c.pad = 0;
                bad |= q[1] != 0;
            }
            out.status = bad ? PA_ELEM_NUMERIC_FAIL : PA_ELEM_OK;
            s->finished.push_back(out);
            L.retired[i] = 1;
            L.n_retired++;
            s->in_flight--;
            s->st.finished++;
            if (n_finished) ++*n_finished;
        }
        if (!ended) continue;
        if (L.n_retired != nc) return fail(e, PA_E_DEVICE, "a serving launch ended with elements not marked complete");
        for (size_t i = 0; i < L.lbatch.size(); i++) {
            const LRequest &r = L.lbatch[i];
            pa_serve_result out;
            memset(&out, 0, sizeof out);
            out.ticket = r.ticket; out.element = r.eid; out.kind = PA_SERVE_LOGLIK;
            out.loglik = L.h_lout[i];
            out.status = std::isfinite(out.loglik) ? PA_ELEM_OK : PA_ELEM_NUMERIC_FAIL;
            out.arrival_s = r.arrival; out.admit_s = r.admit; out.finish_s = now;
            s->finished.push_back(out);
            s->in_flight--;
            s->st.finished++;
            if (n_finished) ++*n_finished;
        }
        float ms = 0;
        HIPCHK(e, hipEventElapsedTime(&ms, L.ev0, L.ev1));
        s->st.device_ms += ms;
        s->pool.release(L.pool_off, L.pool_len);
        L.pool_off = L.pool_len = 0;
        L.batch.clear();
        L.lbatch.clear();
        L.busy = false;
    }
    return PA_OK;
}

// Fills free lanes while requests wait.  A lane is started when no launch is
// in flight, or when fewer elements are in flight than the chain kernel has
// resident workgroups, i.e. when retiring elements have freed workgroup
// slots; the arenas are uploaded first if elements were added (after every
// launch that reads the old ones has ended).
static pa_status serve_admit(pa_server *s)
{
    pa_engine *e = s->e;
    if (s->waiting.empty() && s->lwait.empty()) return PA_OK;
    pa_status st;
    if (e->dev_dirty) {
        if ((st = serve_collect(s, true, nullptr)) != PA_OK) return st;   // no lane may read the old arenas
        if ((st = upload(e)) != PA_OK) return st;
    }
    const int gmax_c = chain_grid_max(e, s->fp64, s->BS, s->shape);
    for (Lane &L : s->lanes) {
        if (L.busy) continue;
        if (s->waiting.empty() && s->lwait.empty()) break;
        bool busy = false;
        for (const Lane &o : s->lanes) busy |= o.busy;
        if (busy && s->in_flight >= gmax_c) break;
        const int64_t before = s->st.launches;
        if ((st = lane_launch(s, L, busy)) != PA_OK) return st;
        if (s->st.launches == before) break;   // nothing fitted the free pool
    }
    return PA_OK;
}

static bool serve_busy(const pa_server *s)
{
    for (const Lane &L : s->lanes)
        if (L.busy) return true;
    return false;
}

extern "C" pa_status pa_serve_step(pa_server *s, int32_t wait, int32_t *n_finished)
{
    if (!s) return PA_E_INPUT;
    pa_engine *e = s->e;
    if (n_finished) *n_finished = 0;
    HIPCHK(e, hipSetDevice(e->device));
    pa_status st;
    if ((st = serve_collect(s, false, n_finished)) != PA_OK) return st;
    if ((st = serve_admit(s)) != PA_OK) return st;
    if (!wait) return PA_OK;
    // wait = 1: run until nothing is in flight and nothing admissible waits
    while (serve_busy(s)) {
        if ((st = serve_collect(s, false, n_finished)) != PA_OK) return st;
        if ((st = serve_admit(s)) != PA_OK) return st;
        if (serve_busy(s)) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    return PA_OK;
}

extern "C" pa_status pa_serve_poll(pa_server *s, pa_serve_result *out, int32_t max, int32_t *n)
{
    if (!s || !n || (max > 0 && !out)) return PA_E_INPUT;
    int k = 0;
    while (k < max && !s->finished.empty()) {
        out[k++] = s->finished.front();
        s->finished.pop_front();
    }
    *n = k;
    return PA_OK;
}

extern "C" pa_status pa_serve_get_stats(pa_server *s, pa_serve_stats *out)
{
    if (!s || !out) return PA_E_INPUT;
    *out = s->st;
    out->waiting = (int32_t)(s->waiting.size() + s->lwait.size());
    out->in_flight = s->in_flight;
    out->scratch_bytes_in_use = (double)s->pool.used;
    int busy = 0;
    for (const Lane &L : s->lanes) busy += L.busy;
    out->lanes_busy = busy;
    return PA_OK;
}

extern "C" void pa_serve_destroy(pa_server *s)
{
    if (!s) return;
    (void)hipSetDevice(s->e->device);
    serve_free(s);
    delete s;
}
