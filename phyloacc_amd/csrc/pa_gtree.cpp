// pa_gtree.cpp -- PhyloAcc-GT's gene trees inside the species tree: the
// embedding of a gene tree held as topology + heights, its multispecies-
// coalescent prior, the missing-gene term, the substitution length of every
// gene branch, and the batched node-height move whose proposals are scored in
// one likelihood launch.
//
// GTree (src/PhyloAcc-GT/genetree.hpp:16-127) caches, per species branch, the
// gene lineages that cross it (parent_gene2) and the gene nodes born in it
// (temp_coal), and updates those caches move by move.  Here they are derived
// from the heights whenever they are needed: a gene node lies in the lowest
// species, on the way up from its first leaf, whose parent species is not
// lower than the node, and a gene lineage crosses every species from its own
// node's up to its parent's.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <memory>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/phyloacc_gpu.h"
#include "pa_common.h"
#include "pa_gtcore.h"
#include "pa_gtdev.h"
#include "pa_kernels.h"
#include "pa_internal.h"
#include "pa_rng.h"

using namespace pa;

namespace {

struct Sp {
    int S, N;
    const int32_t *up;     // parent species, N above the root (the root is N - 1)
    const double *H, *theta;
    std::vector<int32_t> post;          // species in post-order
    std::vector<int32_t> kid0, kid1;    // children of each internal species
    double top(int s) const { return up[s] < N ? H[up[s]] : INFINITY; }
    gt::SpView view() const { return {S, N, up, H, theta, post.data(), kid0.data(), kid1.data()}; }
};

// One gene tree with its embedding.  where[v]: species branch holding gene
// node v; end[v]: the species where lineage v ends (its parent's, or the
// species root for the gene root).
struct Embed {
    const Sp *sp;
    int N, root = -1;
    const int32_t *ch, *par;
    const double *h;
    std::vector<int32_t> where, end, order;

    gt::TreeView view()
    {
        return {sp->S, N, root, ch, par, order.data(), h, where.data(), end.data()};
    }
    gt::TreeView view() const
    {
        return {sp->S, N, root, ch, par, order.data(), h, const_cast<int32_t *>(where.data()),
                const_cast<int32_t *>(end.data())};
    }

    bool build(const Sp &s, const int32_t *children, const int32_t *parent, const double *height)
    {
        sp = &s; N = s.N; ch = children; par = parent; h = height;
        where.assign(N, -1);
        end.assign(N, -1);
        order.clear();
        for (int v = 0; v < N; v++)
            if (par[v] == N) { if (root >= 0) return false; root = v; }
        if (root < 0) return false;
        // post-order from the root; a node's species continues its first child's climb
        thread_local std::vector<uint8_t> seen;
        thread_local std::vector<std::pair<int, int>> st;
        seen.assign(N, 0);
        st.assign(1, {root, 0});
        seen[root] = 1;
        int leaves = 0;
        while (!st.empty()) {
            auto &[v, k] = st.back();
            if (v >= s.S && k < 2) {
                const int c = ch[2 * v + k];
                k++;
                if (c < 0 || c >= N || par[c] != v || seen[c] || !(h[c] <= h[v]) || !std::isfinite(h[v])) return false;
                seen[c] = 1;
                st.push_back({c, 0});
                continue;
            }
            leaves += v < s.S;
            order.push_back(v);
            st.pop_back();
        }
        if (leaves != s.S || (int)order.size() != N) return false;
        gt::TreeView g = view();
        gt::embed(s.view(), g);
        return true;
    }
    // gene node v's height has changed in place (children and parent
    // unchanged, the new height between them)
    void repoint(int v)
    {
        gt::TreeView g = view();
        gt::repoint(sp->view(), g, v);
    }
    bool internal(int v) const { return v >= sp->S; }
    // lineage v crosses species x
    bool crosses(int v, int x) const { return gt::crosses(sp->view(), view(), v, x); }
};

double prior_of(const Embed &g)
{
    thread_local std::vector<int32_t> cnt, off, leave;
    thread_local std::vector<double> hs;
    cnt.resize(g.N); off.resize(g.N + 1); leave.resize(g.N); hs.resize(g.N);
    return gt::prior(g.sp->view(), g.view(), cnt.data(), off.data(), leave.data(), hs.data());
}

void missing_of(const Embed &g, const uint8_t *leaf_missing, std::vector<uint8_t> &miss)
{
    miss.assign(g.N, 0);
    gt::missing(g.view(), leaf_missing, miss.data());
}

double logpz_of(const Embed &g, const std::vector<uint8_t> &miss, const int32_t *Z, double c2m, double n2m)
{
    return gt::logpz(g.sp->view(), g.view(), miss.data(), Z, c2m, n2m);
}

void lengths_of(const Embed &g, const int32_t *Z, double r1, double r2, double *len)
{
    gt::lengths(g.sp->view(), g.view(), Z, r1, r2, len);
}

// Runs fn(i) for i < n on up to 16 host threads (one below 256 elements);
// fn must write only what belongs to element i.
template <typename F>
void for_elements(int n, F &&fn)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 256) nt = 1;
    if (nt == 1) {
        for (int i = 0; i < n; i++) fn(i);
        return;
    }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
            for (int i = (int)t; i < n; i += (int)nt) fn(i);
        });
    for (auto &x : th) x.join();
}

// The trees must be the elements' stored trees: checked on the element
// threads; on a mismatch the first failing element (in batch order) is
// checked again alone, so the recorded message is that element's.  lm
// receives the leaves' missing flags [n][S].
pa_status stored_trees(pa_engine *e, const int32_t *ids, int n, int S, int N, const int32_t *gchildren, uint8_t *lm)
{
    std::vector<pa_status> st(n, PA_OK);
    for_elements(n, [&](int i) { st[i] = element_gene_tree(e, ids[i], lm + (size_t)i * S, gchildren + (size_t)i * 2 * N); });
    for (int i = 0; i < n; i++)
        if (st[i] != PA_OK) return element_gene_tree(e, ids[i], lm + (size_t)i * S, gchildren + (size_t)i * 2 * N);
    return PA_OK;
}

pa_status check_sp(pa_engine *e, const pa_species_tree *t, Sp &s)
{
    if (!t || !t->parent || !t->height || !t->theta) return engine_fail(e, PA_E_INPUT, "null species tree");
    s.S = t->S; s.N = t->N; s.up = t->parent; s.H = t->height; s.theta = t->theta;
    if (s.S < 2 || s.N != 2 * s.S - 1) return engine_fail(e, PA_E_INPUT, "species tree must have N = 2S - 1 nodes");
    int roots = 0;
    for (int x = 0; x < s.N; x++) {
        if (s.up[x] == s.N) roots++;
        else if (s.up[x] < s.S || s.up[x] > s.N || !(s.H[s.up[x]] >= s.H[x]))
            return engine_fail(e, PA_E_INPUT, "species node " + std::to_string(x) + ": bad parent or height");
        if (x >= s.S && !(s.theta[x] > 0)) return engine_fail(e, PA_E_INPUT, "theta must be positive on internal species");
    }
    if (roots != 1 || s.up[s.N - 1] != s.N) return engine_fail(e, PA_E_INPUT, "species tree must have one root, N - 1");
    s.kid0.assign(s.N, -1);
    s.kid1.assign(s.N, -1);
    for (int x = 0; x < s.N - 1; x++) {
        const int p = s.up[x];
        if (s.kid0[p] < 0) s.kid0[p] = x;
        else if (s.kid1[p] < 0) s.kid1[p] = x;
        else return engine_fail(e, PA_E_INPUT, "species tree is not binary");
    }
    for (int x = s.S; x < s.N; x++)
        if (s.kid1[x] < 0) return engine_fail(e, PA_E_INPUT, "species tree is not binary");
    for (int x = 0; x < s.S; x++)
        if (s.kid0[x] >= 0) return engine_fail(e, PA_E_INPUT, "species leaves must be 0..S-1");
    s.post.clear();
    std::vector<std::pair<int, int>> st{{s.N - 1, 0}};
    while (!st.empty()) {
        const int x = st.back().first, k = st.back().second;
        if (x >= s.S && k < 2) {
            st.back().second++;
            st.push_back({k ? s.kid1[x] : s.kid0[x], 0});
            continue;
        }
        s.post.push_back(x);
        st.pop_back();
    }
    if ((int)s.post.size() != s.N) return engine_fail(e, PA_E_INPUT, "species tree does not reach every node");
    return PA_OK;
}

const char *BAD_GENE_TREE = "gene tree is not a binary tree over the species leaves";

}  // namespace

extern "C" pa_status pa_gt_prior(const pa_species_tree *t, const int32_t *children, const int32_t *parent,
                                 const double *height, double *out)
{
    Sp s;
    if (!children || !parent || !height || !out || check_sp(nullptr, t, s) != PA_OK) return PA_E_INPUT;
    Embed g;
    if (!g.build(s, children, parent, height)) return PA_E_INPUT;
    *out = prior_of(g);
    return PA_OK;
}

extern "C" pa_status pa_gt_logpz(const pa_species_tree *t, const int32_t *children, const int32_t *parent,
                                 const double *height, const int32_t *leaf_missing, const int32_t *Z,
                                 double cons_to_mis, double ncons_to_mis, double *out)
{
    Sp s;
    if (!children || !parent || !height || !leaf_missing || !Z || !out || check_sp(nullptr, t, s) != PA_OK) return PA_E_INPUT;
    Embed g;
    if (!g.build(s, children, parent, height)) return PA_E_INPUT;
    std::vector<uint8_t> lm(s.S), miss;
    for (int i = 0; i < s.S; i++) lm[i] = leaf_missing[i] != 0;
    missing_of(g, lm.data(), miss);
    *out = logpz_of(g, miss, Z, cons_to_mis, ncons_to_mis);
    return PA_OK;
}

extern "C" pa_status pa_gt_lengths(const pa_species_tree *t, const int32_t *children, const int32_t *parent,
                                   const double *height, const int32_t *Z, double r1, double r2, double *lengths)
{
    Sp s;
    if (!children || !parent || !height || !Z || !lengths || check_sp(nullptr, t, s) != PA_OK) return PA_E_INPUT;
    Embed g;
    if (!g.build(s, children, parent, height)) return PA_E_INPUT;
    lengths_of(g, Z, r1, r2, lengths);
    return PA_OK;
}

namespace {

using HeightMove = gt::Window;

HeightMove height_window(const Embed &g, int v, double delta) { return gt::window(g.sp->view(), g.view(), v, delta); }

}  // namespace

// Batched GTree::Sample_BranchLen (src/PhyloAcc-GT/genetree.cpp:1658-1965).
extern "C" pa_status pa_gt_branch_len(pa_engine *e, const pa_species_tree *t, const int32_t *ids, int32_t n,
                                      const int32_t *gchildren, const int32_t *gparent, double *gheight,
                                      const int32_t *Z, const double *rates, const double *piA,
                                      const int32_t *gnode, const int32_t *indicator, const int32_t *lens,
                                      double delta, double cons_to_mis, double ncons_to_mis, int32_t m, int32_t slot,
                                      int32_t precision, double *loglik, double *logpz, double *prior,
                                      int32_t *accepted)
{
    if (!e) return PA_E_INPUT;
    if (n < 0 || (n > 0 && (!ids || !gchildren || !gparent || !gheight || !Z || !rates || !gnode || !indicator ||
                            !lens || !loglik || !logpz || !prior || !accepted)))
        return engine_fail(e, PA_E_INPUT, "null argument");
    if (!(delta > 0)) return engine_fail(e, PA_E_INPUT, "delta must be positive");
    Sp s;
    pa_status st = check_sp(e, t, s);
    if (st != PA_OK) return st;
    EngineView ev;
    engine_view(e, &ev);
    if (s.S != ev.S || s.N != ev.N) return engine_fail(e, PA_E_INPUT, "species tree size differs from the model's");
    if (n == 0) return PA_OK;
    const int N = s.N;
    std::vector<uint8_t> lm((size_t)n * s.S);
    if ((st = stored_trees(e, ids, n, s.S, N, gchildren, lm.data())) != PA_OK) return st;
    for (int i = 0; i < n; i++) {
        if (gnode[i] < s.S || gnode[i] >= N) return engine_fail(e, PA_E_INPUT, "gnode must be an internal gene node");
        if (indicator[i] != 0 && indicator[i] != 1) return engine_fail(e, PA_E_INPUT, "indicator must be 0 or 1");
        if (lens[2 * i + 1] < 0 || lens[2 * i + indicator[i]] < 0) return engine_fail(e, PA_E_INPUT, "negative prefix length");
    }
    std::vector<HeightMove> win(n);
    std::vector<double> prop(n, 0.0), rdraw(n, 0.0), prior_new(n, 0.0), lz_new(n, 0.0);
    std::vector<double> rlen((size_t)n * N, 0.0);
    std::vector<int> bad(n, 0);
    const Key key{ev.key0, ev.key1};
    const uint32_t c3 = tagword(0, TAG_GT_BRANCH), c2 = (uint32_t)m;
    // per element, with one reused embedding per host thread: the window and
    // draws, then, with the proposed height written into the caller's row for
    // the moment, the proposal's branch lengths, prior and missing-gene term
    for_elements(n, [&](int i) {
        accepted[i] = 0;
        thread_local Embed g;
        thread_local std::vector<uint8_t> miss;
        const int32_t *ch = gchildren + (size_t)i * 2 * N, *pa = gparent + (size_t)i * N;
        double *h = gheight + (size_t)i * N;
        g.root = -1;
        if (!g.build(s, ch, pa, h)) { bad[i] = 1; return; }
        win[i] = height_window(g, gnode[i], delta);
        if (win[i].status < 0) { bad[i] = 2; return; }
        if (win[i].status != 0) return;
        const double u = slot_uniform(key, (uint32_t)slot, c2, c3, 0);
        prop[i] = std::max(u, 1e-5) * (win[i].upper - win[i].lower) + win[i].lower;
        rdraw[i] = slot_uniform(key, (uint32_t)slot, c2, c3, 1);
        const double old = h[gnode[i]];
        h[gnode[i]] = prop[i];
        g.repoint(gnode[i]);
        lengths_of(g, Z + (size_t)i * N, rates[2 * i], rates[2 * i + 1], &rlen[(size_t)i * N]);
        prior_new[i] = prior_of(g);
        missing_of(g, &lm[(size_t)i * s.S], miss);
        lz_new[i] = logpz_of(g, miss, Z + (size_t)i * N, cons_to_mis, ncons_to_mis);
        h[gnode[i]] = old;
    });
    for (int i = 0; i < n; i++) {
        if (bad[i] == 1) return engine_fail(e, PA_E_INPUT, std::string(BAD_GENE_TREE) + " (element " + std::to_string(i) + ")");
        if (bad[i] == 2)
            return engine_fail(e, PA_E_INPUT, "element " + std::to_string(i) + ": no species holds both children of gene node " +
                                                  std::to_string(gnode[i]));
    }
    // the proposals' prefix likelihoods, one launch
    std::vector<int32_t> run, rid, l0, l1;
    std::vector<double> rl, rpi;
    for (int i = 0; i < n; i++) {
        if (win[i].status != 0) continue;
        run.push_back(i);
        rid.push_back(ids[i]);
        const int ind = indicator[i];
        l0.push_back(lens[2 * i + ind]);
        l1.push_back(std::min(lens[2 * i + 1], lens[2 * i + ind]));
        if (piA) rpi.push_back(piA[i]);
    }
    const int k = (int)run.size();
    if (k == 0) return PA_OK;
    const double *lenp = rlen.data();
    if (k < n) {
        rl.resize((size_t)k * N);
        for (int j = 0; j < k; j++) memcpy(&rl[(size_t)j * N], &rlen[(size_t)run[j] * N], sizeof(double) * N);
        lenp = rl.data();
    }
    std::vector<double> o0(k), o1(k);
    if ((st = pa_loglik_lengths_split(e, rid.data(), k, lenp, piA ? rpi.data() : nullptr, l0.data(), l1.data(),
                                      precision, o0.data(), o1.data())) != PA_OK)
        return st;
    for (int j = 0; j < k; j++) {
        const int i = run[j], ind = indicator[i];
        const double lnew[2] = {o0[j], o1[j]};
        const double d = gt::height_ratio(win[i], prop[i], ind, lnew, loglik + 2 * i, logpz + 2 * i, lz_new[i],
                                          prior_new[i], prior[i]);
        if (std::log(rdraw[i]) < d) {
            accepted[i] = 1;
            gheight[(size_t)i * N + gnode[i]] = prop[i];
            loglik[2 * i] = lnew[0];
            loglik[2 * i + 1] = lnew[1];
            if (logpz[2 * i + 1] == 0) logpz[2 * i] = lz_new[i];
            else logpz[2 * i] = logpz[2 * i + 1] = lz_new[i];
            prior[i] = prior_new[i];
        }
    }
    return PA_OK;
}

namespace {

// Inverse-CDF draw over k non-negative weights: the first i with u * total <
// p[0] + ... + p[i], or the last positive-weight index when rounding puts
// the target past the end.
int categorical_n(const double *p, int k, double u)
{
    double tot = 0;
    int last = 0;
    for (int i = 0; i < k; i++) {
        tot += p[i];
        if (p[i] > 0) last = i;
    }
    const double target = u * tot;
    double cum = 0;
    for (int i = 0; i < k - 1; i++) {
        cum += p[i];
        if (target < cum) return i;
    }
    return last;
}

// Lineage v of a tree whose node species are `where`: from where[v] up to the
// species of its parent gene node (pn), or to the species root when pn is N.
bool lineage_crosses(const Sp &s, const std::vector<int> &where, int v, int pn, int x)
{
    const int top = pn < s.N ? where[pn] : s.N - 1;
    for (int y = where[v];; y = s.up[y]) {
        if (y == x) return true;
        if (y == top || s.up[y] >= s.N) return false;
    }
}

struct TreeMove {
    // candidates: node (N for above the gene root), kind (0 node, 1 root pi-average, 2 root first option,
    // 3 the first candidate when it is the root: score 0), substitution distance
    std::vector<int> cnode, ckind, csp;
    std::vector<double> clen;
    int bp = -1, cb = 0, bs = -1, bpp = -1, ss = -1;
    double curG = 0, curN = 0, cur_unif = 0;
    std::vector<int32_t> ch2, par2;        // the pruned tree
    int root2 = -1;
};

// GTree::Sample_tree2's removal and candidate walk (src/PhyloAcc-GT/
// genetree.cpp:2147-2380, Remove_branch :737-808), unconstrained candidates.
void tree_candidates(const Embed &g, int branch, int root_ingrp, const int32_t *Z, double r1, double r2, TreeMove &t)
{
    const Sp &s = *g.sp;
    const int N = s.N;
    t.curG = prior_of(g);
    t.ss = g.where[branch];
    t.bp = g.par[branch];
    t.cb = g.ch[2 * t.bp] == branch ? 0 : 1;
    t.bs = g.ch[2 * t.bp + 1 - t.cb];
    t.bpp = g.par[t.bp];
    const double thr = s.theta[N - 1];
    t.curN = 10 * thr / 2;
    if (t.curN < g.h[g.root]) t.curN = g.h[g.root] + 5 * thr / 2;
    double lowc = 0;
    for (int x = t.ss; x < N; x = s.up[x])
        if (g.crosses(t.bs, x)) { lowc = std::max(s.H[x], std::max(g.h[branch], g.h[t.bs])); break; }
    t.cur_unif = -std::log((t.bpp < N ? g.h[t.bpp] : t.curN) - lowc);
    // Remove_branch: the sibling takes bp's place
    t.ch2.assign(g.ch, g.ch + 2 * N);
    t.par2.assign(g.par, g.par + N);
    t.par2[t.bs] = t.bpp;
    t.root2 = g.root;
    if (t.bpp < N) t.ch2[2 * t.bpp + (g.ch[2 * t.bpp] == t.bp ? 0 : 1)] = t.bs;
    else t.root2 = t.bs;
    t.par2[t.bp] = -1;
    // gene nodes born in each species, by height (bp is gone; branch's own
    // subtree stays, below branch, and is passed over)
    // (per-species ranges of one list: born[boff[x] .. boff[x + 1]))
    thread_local std::vector<int> bcnt, boff, born;
    bcnt.assign(N, 0);
    for (int v : g.order)
        if (g.internal(v) && v != t.bp) bcnt[g.where[v]]++;
    boff.resize(N + 1);
    boff[0] = 0;
    for (int x = 0; x < N; x++) boff[x + 1] = boff[x] + bcnt[x];
    born.resize(boff[N]);
    for (int x = 0; x < N; x++) bcnt[x] = boff[x];
    for (int v : g.order)
        if (g.internal(v) && v != t.bp) born[bcnt[g.where[v]]++] = v;
    for (int x = 0; x < N; x++)
        if (boff[x + 1] - boff[x] > 1)
            std::sort(born.begin() + boff[x], born.begin() + boff[x + 1], [&](int a, int c) { return g.h[a] < g.h[c]; });
    auto rate = [&](int x) { return Z[x] == 2 ? r2 : (Z[x] == 1 ? r1 : 1.0); };
    auto push = [&](int node, int kind, int x, double len) {
        t.cnode.push_back(node); t.ckind.push_back(kind); t.csp.push_back(x); t.clen.push_back(len);
    };
    double cur = 0;
    for (int x = t.ss; x < N;) {
        const double base = std::max(s.H[x], g.h[branch]);
        const int up = s.up[x];
        if (boff[x] == boff[x + 1]) {
            if (up == N) {
                push(N, t.cnode.empty() ? 3 : 2, x, rate(x) * (t.curN - base) + cur);
                break;
            }
            cur += (s.H[up] - base) * rate(x);
            x = up;
            continue;
        }
        for (int j = boff[x]; j < boff[x + 1]; j++) {
            const int v = born[j];
            if (x == t.ss && g.h[v] <= g.h[branch]) continue;
            push(v, 0, x, rate(x) * (g.h[v] - base) + cur);
        }
        if (up == N) {
            const double ht = t.cnode.empty() ? base : g.h[t.cnode.back()];
            push(N, 1, x, cur + rate(x) * (t.curN - ht));
        }
        if (up < N) cur += (s.H[up] - base) * rate(x);
        if (x == root_ingrp && !t.cnode.empty()) break;
        x = up;
    }
}

}  // namespace

// Batched GTree::Update_Tg with tosample = true: the histories stay in the
// engine's device arena (rows packed, node-major) for the next regraft; Tg,
// when given, also receives them in the caller's [n][N][stride] layout.
extern "C" pa_status pa_gt_sample_history(pa_engine *e, const pa_species_tree *t, const int32_t *ids, int32_t n,
                                          const int32_t *gchildren, const int32_t *gparent, const double *gheight,
                                          const int32_t *Z, const double *rates, const double *piA,
                                          const int32_t *len, int32_t m, int32_t precision, int8_t *Tg,
                                          int32_t stride)
{
    if (!e) return PA_E_INPUT;
    if (n < 0 || (n > 0 && (!ids || !gchildren || !gparent || !gheight || !Z || !rates || !len)))
        return engine_fail(e, PA_E_INPUT, "null argument");
    if (Tg && stride <= 0) return engine_fail(e, PA_E_INPUT, "stride must be positive");
    Sp s;
    pa_status st = check_sp(e, t, s);
    if (st != PA_OK) return st;
    EngineView ev;
    engine_view(e, &ev);
    if (s.S != ev.S || s.N != ev.N) return engine_fail(e, PA_E_INPUT, "species tree size differs from the model's");
    if (n == 0) return PA_OK;
    const int N = s.N;
    std::unique_ptr<double[]> lens(new double[(size_t)n * N]);   // every entry written below
    std::vector<int64_t> off(n);
    std::vector<int32_t> ns(n);
    int64_t total = 0;
    {
        std::vector<uint8_t> lm((size_t)n * s.S);
        if ((st = stored_trees(e, ids, n, s.S, N, gchildren, lm.data())) != PA_OK) return st;
    }
    for (int i = 0; i < n; i++) {
        if (len[i] < 0 || (Tg && len[i] > stride)) return engine_fail(e, PA_E_INPUT, "history length negative or beyond stride");
        off[i] = total;
        ns[i] = len[i];
        total += (int64_t)N * len[i];
    }
    std::vector<int> bad(n, 0);
    for_elements(n, [&](int i) {
        thread_local Embed g;   // reused by the thread's elements; build expects no root yet
        g.root = -1;
        if (!g.build(s, gchildren + (size_t)i * 2 * N, gparent + (size_t)i * N, gheight + (size_t)i * N)) { bad[i] = 1; return; }
        lengths_of(g, Z + (size_t)i * N, rates[2 * i], rates[2 * i + 1], &lens[(size_t)i * N]);
    });
    for (int i = 0; i < n; i++)
        if (bad[i]) return engine_fail(e, PA_E_INPUT, std::string(BAD_GENE_TREE) + " (element " + std::to_string(i) + ")");
    GtHistory &H = engine_history(e);
    H.valid = false;
    if ((st = engine_history_reserve(e, (size_t)std::max<int64_t>(total, 8))) != PA_OK) return st;
    GtExtra x;
    x.d_hist = H.d;
    x.hist_off = off.data();
    x.hist_ns = ns.data();
    x.hist_len = len;
    x.hist_draw = 1;
    x.hist_sweep = m;
    std::vector<double> o(n), o1(n);
    if ((st = gt_launch(e, ids, n, lens.get(), piA, nullptr, nullptr, precision, o.data(), o1.data(), x)) != PA_OK)
        return st;
    H.ids.assign(ids, ids + n);
    H.len.assign(len, len + n);
    H.off = off;
    H.valid = true;
    if (Tg) {
        std::vector<int8_t> host((size_t)std::max<int64_t>(total, 1));
        if (hipMemcpy(host.data(), H.d, (size_t)total, hipMemcpyDeviceToHost) != hipSuccess)
            return engine_fail(e, PA_E_DEVICE, "history copy failed");
        for (int i = 0; i < n; i++)
            for (int v = 0; v < N; v++)
                memcpy(Tg + ((size_t)i * N + v) * stride, &host[(size_t)off[i] + (size_t)v * len[i]], (size_t)len[i]);
    }
    return PA_OK;
}

// Batched GTree::Sample_tree2 (src/PhyloAcc-GT/genetree.cpp:2138-2518).
extern "C" pa_status pa_gt_sample_tree(pa_engine *e, const pa_species_tree *t, const int32_t *ids, int32_t n,
                                       int32_t *gchildren, int32_t *gparent, double *gheight, const int32_t *Z,
                                       const double *rates, const double *piA, const int32_t *branch,
                                       const int32_t *indicator, const int32_t *lens, const int8_t *Tg,
                                       int32_t stride, double cons_to_mis, double ncons_to_mis, int32_t root_ingrp,
                                       int32_t m, int32_t slot, int32_t precision, double *loglik, double *logpz,
                                       int32_t *accepted)
{
    if (!e) return PA_E_INPUT;
    if (n < 0 || (n > 0 && (!ids || !gchildren || !gparent || !gheight || !Z || !rates || !branch || !indicator ||
                            !lens || !loglik || !logpz || !accepted)))
        return engine_fail(e, PA_E_INPUT, "null argument");
    Sp s;
    pa_status st = check_sp(e, t, s);
    if (st != PA_OK) return st;
    EngineView ev;
    engine_view(e, &ev);
    if (s.S != ev.S || s.N != ev.N) return engine_fail(e, PA_E_INPUT, "species tree size differs from the model's");
    if (root_ingrp < s.S || root_ingrp >= s.N) return engine_fail(e, PA_E_INPUT, "root_ingrp must be an internal species");
    if (n == 0) return PA_OK;
    {
        std::vector<int32_t> u(ids, ids + n);
        std::sort(u.begin(), u.end());
        if (std::adjacent_find(u.begin(), u.end()) != u.end())
            return engine_fail(e, PA_E_INPUT, "an element appears twice in one tree move");
    }
    const int N = s.N;
    const Key key{ev.key0, ev.key1};
    const uint32_t c3 = tagword(0, TAG_GT_TREE), c2 = (uint32_t)m;
    std::vector<Embed> emb(n);
    std::vector<TreeMove> mv(n);
    std::vector<uint8_t> lm((size_t)n * s.S);
    std::unique_ptr<double[]> cur_len(new double[(size_t)n * N]);   // every entry written below
    if ((st = stored_trees(e, ids, n, s.S, N, gchildren, lm.data())) != PA_OK) return st;
    for (int i = 0; i < n; i++) {
        const int ind = indicator[i];
        if (ind != 0 && ind != 1) return engine_fail(e, PA_E_INPUT, "indicator must be 0 or 1");
        if (lens[2 * i + ind] < 0 || lens[2 * i + 1] < 0 || (Tg && lens[2 * i + ind] > stride))
            return engine_fail(e, PA_E_INPUT, "prefix length negative or beyond the history's stride");
        if (branch[i] < 0 || branch[i] >= N || gparent[(size_t)i * N + branch[i]] == N)
            return engine_fail(e, PA_E_INPUT, "branch must be a gene node other than the root");
    }
    std::vector<int> bad(n, 0);
    for_elements(n, [&](int i) {
        if (!emb[i].build(s, gchildren + (size_t)i * 2 * N, gparent + (size_t)i * N, gheight + (size_t)i * N)) {
            bad[i] = 1;
            return;
        }
        tree_candidates(emb[i], branch[i], root_ingrp, Z + (size_t)i * N, rates[2 * i], rates[2 * i + 1], mv[i]);
        lengths_of(emb[i], Z + (size_t)i * N, rates[2 * i], rates[2 * i + 1], &cur_len[(size_t)i * N]);
    });
    int cmax = 1;
    for (int i = 0; i < n; i++) {
        if (bad[i]) return engine_fail(e, PA_E_INPUT, std::string(BAD_GENE_TREE) + " (element " + std::to_string(i) + ")");
        cmax = std::max(cmax, (int)mv[i].cnode.size());
    }
    // launch 1: the current trees' partials score every candidate
    std::vector<int32_t> cb(n), cn(n), cnode((size_t)n * cmax, 0), ckind((size_t)n * cmax, 1), csites(n);
    std::vector<double> clen((size_t)n * cmax, 0.0), cscore((size_t)n * cmax, 0.0);
    for (int i = 0; i < n; i++) {
        cb[i] = branch[i];
        cn[i] = (int)mv[i].cnode.size();
        csites[i] = lens[2 * i + indicator[i]];
        for (int q = 0; q < cn[i]; q++) {
            cnode[(size_t)i * cmax + q] = mv[i].cnode[q] < N ? mv[i].cnode[q] : 0;
            ckind[(size_t)i * cmax + q] = mv[i].ckind[q] == 3 ? 1 : mv[i].ckind[q];
            clen[(size_t)i * cmax + q] = mv[i].clen[q];
        }
    }
    {
        // the histories: the engine's last draw (Tg null), or the caller's
        GtExtra x;
        std::vector<int64_t> off(n);
        std::vector<int32_t> ns(n);
        int8_t *dh = nullptr;
        if (!Tg) {
            GtHistory &Hs = engine_history(e);
            if (!Hs.valid || (int)Hs.ids.size() != n || !std::equal(Hs.ids.begin(), Hs.ids.end(), ids))
                return engine_fail(e, PA_E_STATE, "no device history for these elements: call pa_gt_sample_history "
                                                  "with the same ids first, or pass Tg");
            for (int i = 0; i < n; i++) {
                if (Hs.len[i] < csites[i]) return engine_fail(e, PA_E_INPUT, "the device history covers fewer sites");
                off[i] = Hs.off[i];
                ns[i] = Hs.len[i];
            }
            dh = Hs.d;
        } else {
            int64_t total = 0;
            for (int i = 0; i < n; i++) { off[i] = total; ns[i] = csites[i]; total += (int64_t)N * csites[i]; }
            std::vector<int8_t> host((size_t)std::max<int64_t>(total, 1));
            for (int i = 0; i < n; i++)
                for (int v = 0; v < N; v++)
                    memcpy(&host[(size_t)off[i] + (size_t)v * ns[i]], Tg + ((size_t)i * N + v) * stride, (size_t)ns[i]);
            // a separate arena: the engine's history stays as it was
            if (hipMalloc(&dh, host.size()) != hipSuccess || hipMemcpy(dh, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) {
                (void)hipFree(dh);
                return engine_fail(e, PA_E_DEVICE, "history upload failed");
            }
        }
        x.d_hist = dh;
        x.hist_off = off.data();
        x.hist_ns = ns.data();
        x.cand_branch = cb.data(); x.cand_n = cn.data(); x.cand_node = cnode.data(); x.cand_kind = ckind.data();
        x.cand_sites = csites.data(); x.cand_len = clen.data(); x.cand_max = cmax; x.cand_score = cscore.data();
        std::vector<double> o(n), o1(n);
        st = gt_launch(e, ids, n, cur_len.get(), piA, nullptr, nullptr, precision, o.data(), o1.data(), x);
        if (Tg) (void)hipFree(dh);
        if (st != PA_OK) return st;
    }
    // draw each element's regraft; build the proposed trees
    // the proposed trees and candidate probabilities, one row per element
    // (every row written in the draw below before it is read)
    std::unique_ptr<int32_t[]> nch_buf(new int32_t[(size_t)n * 2 * N]), npar_buf(new int32_t[(size_t)n * N]);
    std::unique_ptr<double[]> nh_buf(new double[(size_t)n * N]), sp_buf(new double[(size_t)n * cmax]);
    auto nch = [&](int i) { return &nch_buf[(size_t)i * 2 * N]; };
    auto npar = [&](int i) { return &npar_buf[(size_t)i * N]; };
    auto nh = [&](int i) { return &nh_buf[(size_t)i * N]; };
    auto sprob = [&](int i) { return &sp_buf[(size_t)i * cmax]; };
    std::vector<int> get(n), pos(n);
    std::vector<double> new_unif(n), rdraw(n);
    // every row of tabs is written by gene_table before the launch reads it (a failure returns first)
    std::unique_ptr<int32_t[]> tabs(new int32_t[(size_t)n * tab_words(N)]);
    std::vector<int32_t> hdrs((size_t)n * 4), l0(n), l1(n);
    std::unique_ptr<double[]> nlen(new double[(size_t)n * N]);   // every entry written below
    std::vector<int> gbad(n, 0);
    std::vector<double> prior2(n), lz2(n);   // the proposed trees' prior and missing-gene term
    for_elements(n, [&](int i) {
        TreeMove &t = mv[i];
        const Embed &g = emb[i];
        const int nc = (int)t.cnode.size();
        double *p = sprob(i);
        std::fill(p, p + nc, 1.0);
        get[i] = 0;
        if (nc > 1) {
            double mx = -INFINITY, tot = 0;
            thread_local std::vector<double> sc;
            sc.resize(nc);
            for (int q = 0; q < nc; q++) {
                sc[q] = t.ckind[q] == 3 ? 0.0 : cscore[(size_t)i * cmax + q];
                mx = std::max(mx, sc[q]);
            }
            for (int q = 0; q < nc; q++) { p[q] = std::exp(sc[q] - mx); tot += p[q]; }
            for (int q = 0; q < nc; q++) p[q] /= tot;
            get[i] = categorical_n(p, nc, slot_uniform(key, (uint32_t)slot, c2, c3, 0));
        }
        pos[i] = -1;
        for (int q = 0; q < nc; q++) if (t.cnode[q] == (t.bpp < N ? t.bpp : N)) { pos[i] = q; break; }
        const int nb = t.cnode[get[i]];
        const double thr = s.theta[N - 1];
        int nsib;
        double upp;
        if (nb == N) {
            nsib = t.root2;
            upp = 10 * thr / 2 > g.h[nsib] ? 10 * thr / 2 : g.h[nsib] + 5 * thr / 2;
        } else {
            nsib = t.ch2[2 * nb + (slot_uniform(key, (uint32_t)slot, c2, c3, 1) < 0.5 ? 0 : 1)];
            upp = g.h[nb];
        }
        double low = 0;
        const int stop = s.up[t.csp[get[i]]];
        for (int x = t.ss; x != stop; x = s.up[x])
            if (lineage_crosses(s, g.where, nsib, t.par2[nsib], x)) {
                low = std::max(g.h[branch[i]], std::max(g.h[nsib], s.H[x]));
                break;
            }
        const double u = std::max(slot_uniform(key, (uint32_t)slot, c2, c3, 2), 1e-5);
        new_unif[i] = -std::log(upp - low);
        rdraw[i] = slot_uniform(key, (uint32_t)slot, c2, c3, 3);
        // Graft: bp above nsib, branch keeps its side of bp
        int32_t *ch_i = nch(i), *par_i = npar(i);
        double *h_i = nh(i);
        memcpy(ch_i, t.ch2.data(), sizeof(int32_t) * 2 * N);
        memcpy(par_i, t.par2.data(), sizeof(int32_t) * N);
        memcpy(h_i, g.h, sizeof(double) * N);
        const int bp = t.bp, it2p = par_i[nsib];
        h_i[bp] = u * (upp - low) + low;
        par_i[bp] = it2p;
        par_i[nsib] = bp;
        ch_i[2 * bp + t.cb] = branch[i];
        ch_i[2 * bp + 1 - t.cb] = nsib;
        if (it2p < N) ch_i[2 * it2p + (t.ch2[2 * it2p] == nsib ? 0 : 1)] = bp;
        if (gene_table(e, ids[i], ch_i, &tabs[(size_t)i * tab_words(N)], &hdrs[(size_t)i * 4]) != PA_OK) {
            gbad[i] = 1;
            return;
        }
        thread_local Embed g2;   // the proposed tree's embedding, reused by the thread's elements
        g2.root = -1;
        if (!g2.build(s, ch_i, par_i, h_i)) { gbad[i] = 1; return; }
        lengths_of(g2, Z + (size_t)i * N, rates[2 * i], rates[2 * i + 1], &nlen[(size_t)i * N]);
        // the decision's terms that do not need the likelihoods
        thread_local std::vector<uint8_t> miss;
        missing_of(g2, &lm[(size_t)i * s.S], miss);
        lz2[i] = logpz_of(g2, miss, Z + (size_t)i * N, cons_to_mis, ncons_to_mis);
        prior2[i] = prior_of(g2);
        const int ind = indicator[i];
        l0[i] = lens[2 * i + ind];
        l1[i] = std::min(lens[2 * i + 1], lens[2 * i + ind]);
    });
    for (int i = 0; i < n; i++)
        if (gbad[i]) return engine_fail(e, PA_E_INPUT, "element " + std::to_string(i) + ": regrafted gene tree is not valid");
    // launch 2: the proposed trees' prefix likelihoods
    std::vector<double> o0(n), o1(n);
    {
        GtExtra x;
        x.tabs = tabs.get();
        x.hdrs = hdrs.data();
        if ((st = gt_launch(e, ids, n, nlen.get(), piA, l0.data(), l1.data(), precision, o0.data(), o1.data(), x)) != PA_OK)
            return st;
    }
    for_elements(n, [&](int i) {
        accepted[i] = 0;
        if (pos[i] < 0) return;
        const double lz = lz2[i];
        const int ind = indicator[i];
        const double lnew[2] = {o0[i], o1[i]};
        double d = lnew[ind] - loglik[2 * i + ind];
        if (logpz[2 * i + ind] != 0) d += lz - logpz[2 * i + ind];
        d += prior2[i] - mv[i].curG + mv[i].cur_unif - new_unif[i] + std::log(sprob(i)[pos[i]]) -
             std::log(sprob(i)[get[i]]);
        if (!(std::log(rdraw[i]) < d)) return;
        accepted[i] = 1;
        memcpy(gchildren + (size_t)i * 2 * N, nch(i), sizeof(int32_t) * 2 * N);
        memcpy(gparent + (size_t)i * N, npar(i), sizeof(int32_t) * N);
        memcpy(gheight + (size_t)i * N, nh(i), sizeof(double) * N);
        loglik[2 * i] = lnew[0];
        loglik[2 * i + 1] = lnew[1];
        if (logpz[2 * i + 1] == 0) logpz[2 * i] = lz;
        else logpz[2 * i] = logpz[2 * i + 1] = lz;
    });
    // the per-element work vectors were allocated on the element threads and
    // are freed there too (freed one by one on this thread when the call
    // returned, they took about 50 ms at 20,000 elements of 239 gene nodes,
    // a quarter of the call)
    for_elements(n, [&](int i) {
        emb[i] = Embed();
        mv[i] = TreeMove();
    });
    // the accepted trees become the elements' stored trees
    std::vector<int32_t> sid;
    std::vector<const int32_t *> stab, shdr;
    for (int i = 0; i < n; i++)
        if (accepted[i]) {
            sid.push_back(ids[i]);
            stab.push_back(&tabs[(size_t)i * tab_words(N)]);
            shdr.push_back(&hdrs[(size_t)i * 4]);
        }
    return element_set_trees(e, sid.data(), (int)sid.size(), stab.data(), shdr.data());
}

// ------------------------------------------------ device-resident height moves

struct pa_gt_session {
    pa_engine *e = nullptr;
    int n = 0;
    Sp sp;
    std::vector<int32_t> up_h;
    std::vector<double> H_h, th_h;
    std::vector<int32_t> ids;
    std::vector<double> piA;
    std::vector<void *> bufs;
    // device arrays
    double lc[4] = {0, 0, 0, 0};
    int32_t *up = nullptr, *post = nullptr, *nleaf = nullptr;
    double *H = nullptr, *theta = nullptr, *rate = nullptr, *ldens = nullptr;
    int32_t *ch = nullptr, *par = nullptr, *order = nullptr, *root = nullptr, *Z = nullptr, *where = nullptr;
    double *h = nullptr, *rates = nullptr, *ll = nullptr, *lz = nullptr, *prior = nullptr;
    uint8_t *miss = nullptr;
    int32_t *gnode = nullptr, *ind = nullptr, *acc = nullptr, *wnew = nullptr;
    double *len = nullptr, *prop = nullptr, *rdraw = nullptr, *prior_new = nullptr, *lz_new = nullptr;
    double *out0 = nullptr, *out1 = nullptr;
    gt::Window *win = nullptr;
    // resident partials of the current trees, at one precision
    int res_prec = -1;
    char *big = nullptr;
    uint8_t *alt = nullptr;
    std::vector<int64_t> big_off;
    GtPlan *plan = nullptr;
    std::vector<int32_t> h_gnode, h_ind, h_acc, l0, l1;   // staged for the asynchronous copies
    // sweeps: every move's nodes and decisions on the device, one event pair per move
    int32_t *sw_gnode = nullptr, *sw_acc = nullptr;
    size_t sw_cap = 0;
    std::vector<hipEvent_t> ev;

    template <typename T>
    hipError_t get(T **p, size_t count, const void *src = nullptr)
    {
        void *q = nullptr;
        hipError_t r = hipMalloc(&q, std::max<size_t>(count * sizeof(T), 8));
        if (r != hipSuccess) return r;
        bufs.push_back(q);
        *p = (T *)q;
        if (src && count) r = hipMemcpy(q, src, count * sizeof(T), hipMemcpyHostToDevice);
        return r;
    }
    ~pa_gt_session()
    {
        gt_plan_free(plan);
        for (void *p : bufs) (void)hipFree(p);
        (void)hipFree(big);
        (void)hipFree(sw_gnode);
        (void)hipFree(sw_acc);
        for (hipEvent_t x : ev) (void)hipEventDestroy(x);
    }
};

static void session_prop_args(pa_gt_session *s, GtPropArgs &a)
{
    const int S = s->sp.S, N = s->sp.N;
    a.sp = {S, N, s->up, s->H, s->theta, s->post, nullptr, nullptr};
    a.nleaf = s->nleaf; a.rate = s->rate; a.ldens = s->ldens;
    for (int k = 0; k < 4; k++) a.lc[k] = s->lc[k];
    a.n = s->n; a.S = S; a.N = N;
    a.ch = s->ch; a.par = s->par; a.order = s->order; a.root = s->root; a.where = s->where;
    a.h = s->h; a.miss = s->miss; a.Z = s->Z; a.rates = s->rates; a.gnode = s->gnode;
    a.delta = 1.0; a.key0 = a.key1 = 0; a.m = a.slot = 0; a.lengths_only = 0;
    a.len = s->len; a.win = s->win; a.prop = s->prop; a.rdraw = s->rdraw; a.prior_new = s->prior_new; a.lz_new = s->lz_new;
    a.wnew = s->wnew;
}

// The partials of every element's current tree, kept in HBM at this
// precision (two slots per internal node, all in their first): the current
// trees' lengths, then one full pass that stores every node's partial.
static pa_status session_resident(pa_gt_session *s, int32_t precision)
{
    pa_engine *e = s->e;
    const int n = s->n, S = s->sp.S, N = s->sp.N, fp64 = precision == PA_FP64;
    if (precision != PA_FP32 && precision != PA_FP64) return engine_fail(e, PA_E_INPUT, "precision must be PA_FP32 or PA_FP64");
    hipStream_t stream;
    pa_status st = engine_stream(e, &stream);
    if (st != PA_OK) return st;
    (void)hipStreamSynchronize(stream);
    s->res_prec = -1;
    gt_plan_free(s->plan);
    s->plan = nullptr;
    (void)hipFree(s->big);
    s->big = nullptr;
    s->big_off.assign(n, 0);
    size_t total = 0;
    for (int i = 0; i < n; i++) {
        const int pp = element_ppad(e, s->ids[i]);
        if (pp < 0) return engine_fail(e, PA_E_INPUT, "unknown element id");
        s->big_off[i] = (int64_t)total;
        total += loglik_resident_bytes(N, S, pp, fp64);
    }
    if (hipMalloc(&s->big, std::max<size_t>(total, 256)) != hipSuccess)
        return engine_fail(e, PA_E_OOM, "gene-tree session: no device memory for the resident partials (" +
                                         std::to_string(total >> 20) + " MiB)");
    if (!s->alt && s->get(&s->alt, (size_t)n * N) != hipSuccess) return engine_fail(e, PA_E_OOM, "session slot bits");
    if (hipMemsetAsync(s->alt, 0, (size_t)n * N, stream) != hipSuccess) return engine_fail(e, PA_E_DEVICE, "session slot bits");
    GtPropArgs a;
    session_prop_args(s, a);
    a.lengths_only = 1;
    if (launch_gt_hprop(a, stream) != hipSuccess) return engine_fail(e, PA_E_DEVICE, "session lengths launch failed");
    GtExtra x;
    x.d_len = s->len;
    x.d_row0 = s->out0;
    x.d_row1 = s->out1;
    x.d_big = s->big;
    x.row_big_off = s->big_off.data();
    x.d_alt = s->alt;
    if ((st = gt_plan_build(e, s->ids.data(), n, s->piA.empty() ? nullptr : s->piA.data(), precision, x, &s->plan)) !=
        PA_OK)
        return st;
    if ((st = gt_plan_run(e, s->plan, nullptr, nullptr, nullptr)) != PA_OK || (st = gt_plan_stats(e, s->plan)) != PA_OK)
        return st;
    s->res_prec = precision;
    return PA_OK;
}

extern "C" pa_status pa_gt_session_create(pa_engine *e, const pa_species_tree *t, const int32_t *ids, int32_t n,
                                          const int32_t *gchildren, const int32_t *gparent, const double *gheight,
                                          const int32_t *Z, const double *rates, const double *piA,
                                          const double *loglik, const double *logpz, const double *prior,
                                          double cons_to_mis, double ncons_to_mis, pa_gt_session **out)
{
    if (!e || !out) return PA_E_INPUT;
    *out = nullptr;
    if (n <= 0 || !ids || !gchildren || !gparent || !gheight || !Z || !rates || !loglik || !logpz || !prior)
        return engine_fail(e, PA_E_INPUT, "null argument or empty batch");
    if (!t || !t->parent || !t->height || !t->theta) return engine_fail(e, PA_E_INPUT, "null species tree");
    std::unique_ptr<pa_gt_session> s(new pa_gt_session);
    s->e = e;
    s->n = n;
    // the session keeps its own copy of the species tree
    s->up_h.assign(t->parent, t->parent + t->N);
    s->H_h.assign(t->height, t->height + t->N);
    s->th_h.assign(t->theta, t->theta + t->N);
    pa_species_tree own{t->S, t->N, s->up_h.data(), s->H_h.data(), s->th_h.data()};
    pa_status st = check_sp(e, &own, s->sp);
    if (st != PA_OK) return st;
    EngineView ev;
    engine_view(e, &ev);
    if (s->sp.S != ev.S || s->sp.N != ev.N) return engine_fail(e, PA_E_INPUT, "species tree size differs from the model's");
    const int S = s->sp.S, N = s->sp.N;
    if (gt_hprop_lds(N) > GT_HPROP_LDS_MAX)
        return engine_fail(e, PA_E_INPUT, "species tree too large for the device session (its LDS copy exceeds 64 KiB)");
    std::vector<uint8_t> lm((size_t)n * S);
    if ((st = stored_trees(e, ids, n, S, N, gchildren, lm.data())) != PA_OK) return st;
    // per element: the post-order, root, embedding and missing flags; the
    // trees keep their topology for the session's life, so the flags hold
    // every element's rows are written below before they are uploaded (a bad tree returns first)
    std::unique_ptr<int32_t[]> order(new int32_t[(size_t)n * N]), where(new int32_t[(size_t)n * N]);
    std::unique_ptr<uint8_t[]> miss(new uint8_t[(size_t)n * N]);
    std::vector<int32_t> root(n);
    std::vector<int> bad(n, 0);
    const Sp &sp = s->sp;
    for_elements(n, [&](int i) {
        thread_local Embed g;
        thread_local std::vector<uint8_t> mi;
        g.root = -1;
        if (!g.build(sp, gchildren + (size_t)i * 2 * N, gparent + (size_t)i * N, gheight + (size_t)i * N)) { bad[i] = 1; return; }
        memcpy(&order[(size_t)i * N], g.order.data(), sizeof(int32_t) * N);
        memcpy(&where[(size_t)i * N], g.where.data(), sizeof(int32_t) * N);
        missing_of(g, &lm[(size_t)i * S], mi);
        memcpy(&miss[(size_t)i * N], mi.data(), N);
        root[i] = g.root;
    });
    // per species: the leaves below it, and the prior's 1 / theta and log(2 / theta)
    // with the host's own arithmetic
    std::vector<int32_t> nleaf(N, 0);
    std::vector<double> rate(N), ldens(N);
    for (int q = 0; q < N; q++) {
        const int x = sp.post[q];
        nleaf[x] = x < S ? 1 : nleaf[sp.kid0[x]] + nleaf[sp.kid1[x]];
        rate[x] = 1.0 / sp.theta[x];
        ldens[x] = std::log(2.0 * rate[x]);
    }
    gt::logpz_consts(cons_to_mis, ncons_to_mis, s->lc);
    for (int i = 0; i < n; i++)
        if (bad[i]) return engine_fail(e, PA_E_INPUT, std::string(BAD_GENE_TREE) + " (element " + std::to_string(i) + ")");
    s->ids.assign(ids, ids + n);
    if (piA) {
        for (int i = 0; i < n; i++)
            if (!(piA[i] > 0 && piA[i] < 0.5)) return engine_fail(e, PA_E_INPUT, "piA must lie in (0, 1/2)");
        s->piA.assign(piA, piA + n);
    }
    hipStream_t stream;
    if ((st = engine_stream(e, &stream)) != PA_OK) return st;
    const size_t nN = (size_t)n * N;
    hipError_t r = hipSuccess;
    auto chk = [&](hipError_t x) { if (r == hipSuccess) r = x; };
    chk(s->get(&s->up, N, sp.up)); chk(s->get(&s->H, N, sp.H)); chk(s->get(&s->theta, N, sp.theta));
    chk(s->get(&s->post, N, sp.post.data())); chk(s->get(&s->nleaf, N, nleaf.data()));
    chk(s->get(&s->rate, N, rate.data())); chk(s->get(&s->ldens, N, ldens.data()));
    chk(s->get(&s->where, nN, where.get())); chk(s->get(&s->miss, nN, miss.get())); chk(s->get(&s->wnew, n));
    chk(s->get(&s->ch, 2 * nN, gchildren)); chk(s->get(&s->par, nN, gparent)); chk(s->get(&s->order, nN, order.get()));
    chk(s->get(&s->root, n, root.data())); chk(s->get(&s->h, nN, gheight)); chk(s->get(&s->Z, nN, Z));
    chk(s->get(&s->rates, 2 * (size_t)n, rates));
    chk(s->get(&s->ll, 2 * (size_t)n, loglik)); chk(s->get(&s->lz, 2 * (size_t)n, logpz)); chk(s->get(&s->prior, n, prior));
    chk(s->get(&s->len, nN)); chk(s->get(&s->win, n)); chk(s->get(&s->prop, n)); chk(s->get(&s->rdraw, n));
    chk(s->get(&s->prior_new, n)); chk(s->get(&s->lz_new, n)); chk(s->get(&s->out0, n)); chk(s->get(&s->out1, n));
    chk(s->get(&s->gnode, n)); chk(s->get(&s->ind, n)); chk(s->get(&s->acc, n));
    if (r != hipSuccess) return engine_fail(e, PA_E_OOM, std::string("gene-tree session buffers: ") + hipGetErrorString(r));
    *out = s.release();
    return PA_OK;
}

extern "C" pa_status pa_gt_session_branch_len(pa_gt_session *s, const int32_t *gnode, const int32_t *indicator,
                                              const int32_t *lens, double delta, int32_t m, int32_t slot,
                                              int32_t precision, int32_t *accepted)
{
    if (!s) return PA_E_INPUT;
    pa_engine *e = s->e;
    if (!gnode || !indicator || !lens) return engine_fail(e, PA_E_INPUT, "null argument");
    if (!(delta > 0)) return engine_fail(e, PA_E_INPUT, "delta must be positive");
    const int n = s->n, S = s->sp.S, N = s->sp.N;
    s->l0.resize(n);
    s->l1.resize(n);
    for (int i = 0; i < n; i++) {
        if (gnode[i] < S || gnode[i] >= N) return engine_fail(e, PA_E_INPUT, "gnode must be an internal gene node");
        if (indicator[i] != 0 && indicator[i] != 1) return engine_fail(e, PA_E_INPUT, "indicator must be 0 or 1");
        if (lens[2 * i + 1] < 0 || lens[2 * i + indicator[i]] < 0) return engine_fail(e, PA_E_INPUT, "negative prefix length");
        s->l0[i] = lens[2 * i + indicator[i]];
        s->l1[i] = std::min(lens[2 * i + 1], s->l0[i]);
    }
    pa_status st;
    if (s->res_prec != precision && (st = session_resident(s, precision)) != PA_OK) return st;
    hipStream_t stream;
    if ((st = engine_stream(e, &stream)) != PA_OK) return st;
    s->h_gnode.assign(gnode, gnode + n);
    s->h_ind.assign(indicator, indicator + n);
    if (hipMemcpyAsync(s->gnode, s->h_gnode.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipMemcpyAsync(s->ind, s->h_ind.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, stream) != hipSuccess)
        return engine_fail(e, PA_E_DEVICE, "session upload failed");
    EngineView ev;
    engine_view(e, &ev);
    GtPropArgs a;
    session_prop_args(s, a);
    a.delta = delta;
    a.key0 = ev.key0; a.key1 = ev.key1; a.m = m; a.slot = slot;
    if (launch_gt_hprop(a, stream) != hipSuccess) return engine_fail(e, PA_E_DEVICE, "height proposal launch failed");
    // the proposals' likelihoods: each element's path from the moved node to
    // the root, into the slots its current partials do not use
    if ((st = gt_plan_run(e, s->plan, s->l0.data(), s->l1.data(), s->gnode)) != PA_OK) return st;
    GtDecideArgs d;
    d.n = n; d.N = N;
    d.win = s->win; d.prop = s->prop; d.rdraw = s->rdraw; d.prior_new = s->prior_new; d.lz_new = s->lz_new;
    d.wnew = s->wnew; d.where = s->where;
    d.par = s->par; d.root = s->root; d.alt = s->alt;
    d.lnew0 = s->out0; d.lnew1 = s->out1; d.gnode = s->gnode; d.indicator = s->ind;
    d.h = s->h; d.loglik = s->ll; d.logpz = s->lz; d.prior = s->prior; d.accepted = s->acc;
    if (launch_gt_hdecide(d, stream) != hipSuccess) return engine_fail(e, PA_E_DEVICE, "height decision launch failed");
    if (accepted) {
        s->h_acc.resize(n);
        if (hipMemcpyAsync(s->h_acc.data(), s->acc, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream) != hipSuccess)
            return engine_fail(e, PA_E_DEVICE, "session download failed");
    }
    if (hipStreamSynchronize(stream) != hipSuccess) return engine_fail(e, PA_E_DEVICE, "height moves failed");
    if (accepted) memcpy(accepted, s->h_acc.data(), sizeof(int32_t) * n);
    return gt_plan_stats(e, s->plan);
}

extern "C" pa_status pa_gt_session_sweep(pa_gt_session *s, const int32_t *gnodes, int32_t k, const int32_t *indicator,
                                         const int32_t *lens, double delta, int32_t m, int32_t slot0,
                                         int32_t precision, int32_t *accepted, double *kernel_ms)
{
    if (!s) return PA_E_INPUT;
    pa_engine *e = s->e;
    if (!gnodes || !indicator || !lens || k < 0) return engine_fail(e, PA_E_INPUT, "null argument");
    if (!(delta > 0)) return engine_fail(e, PA_E_INPUT, "delta must be positive");
    if (k == 0) {
        if (kernel_ms) *kernel_ms = 0;
        return PA_OK;
    }
    const int n = s->n, S = s->sp.S, N = s->sp.N;
    const size_t kn = (size_t)k * n;
    for (size_t q = 0; q < kn; q++)
        if (gnodes[q] < S || gnodes[q] >= N) return engine_fail(e, PA_E_INPUT, "gnode must be an internal gene node");
    s->l0.resize(n);
    s->l1.resize(n);
    for (int i = 0; i < n; i++) {
        if (indicator[i] != 0 && indicator[i] != 1) return engine_fail(e, PA_E_INPUT, "indicator must be 0 or 1");
        if (lens[2 * i + 1] < 0 || lens[2 * i + indicator[i]] < 0) return engine_fail(e, PA_E_INPUT, "negative prefix length");
        s->l0[i] = lens[2 * i + indicator[i]];
        s->l1[i] = std::min(lens[2 * i + 1], s->l0[i]);
    }
    pa_status st;
    if (s->res_prec != precision && (st = session_resident(s, precision)) != PA_OK) return st;
    hipStream_t stream;
    if ((st = engine_stream(e, &stream)) != PA_OK) return st;
    if (hipStreamSynchronize(stream) != hipSuccess) return engine_fail(e, PA_E_DEVICE, "session stream failed");
    if (kn > s->sw_cap) {
        (void)hipFree(s->sw_gnode);
        (void)hipFree(s->sw_acc);
        s->sw_gnode = s->sw_acc = nullptr;
        s->sw_cap = 0;
        if (hipMalloc(&s->sw_gnode, kn * 4) != hipSuccess || hipMalloc(&s->sw_acc, kn * 4) != hipSuccess)
            return engine_fail(e, PA_E_OOM, "session sweep buffers");
        s->sw_cap = kn;
    }
    while ((int)s->ev.size() < 2 * k) {
        hipEvent_t x;
        if (hipEventCreate(&x) != hipSuccess) return engine_fail(e, PA_E_DEVICE, "session events");
        s->ev.push_back(x);
    }
    s->h_ind.assign(indicator, indicator + n);
    if (hipMemcpy(s->sw_gnode, gnodes, kn * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->ind, s->h_ind.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice) != hipSuccess)
        return engine_fail(e, PA_E_DEVICE, "session upload failed");
    EngineView ev;
    engine_view(e, &ev);
    // every move on the stream in turn: each reads the state the previous
    // one's decisions left, so nothing waits on the host in between
    for (int j = 0; j < k; j++) {
        int32_t *gn = s->sw_gnode + (size_t)j * n;
        GtPropArgs a;
        session_prop_args(s, a);
        a.gnode = gn;
        a.delta = delta;
        a.key0 = ev.key0; a.key1 = ev.key1; a.m = m; a.slot = slot0 + j;
        if (launch_gt_hprop(a, stream) != hipSuccess) return engine_fail(e, PA_E_DEVICE, "height proposal launch failed");
        if (hipEventRecord(s->ev[2 * j], stream) != hipSuccess) return engine_fail(e, PA_E_DEVICE, "session events");
        if ((st = gt_plan_run(e, s->plan, s->l0.data(), s->l1.data(), gn, j > 0, j > 0)) != PA_OK) return st;
        if (hipEventRecord(s->ev[2 * j + 1], stream) != hipSuccess) return engine_fail(e, PA_E_DEVICE, "session events");
        GtDecideArgs d;
        d.n = n; d.N = N;
        d.win = s->win; d.prop = s->prop; d.rdraw = s->rdraw; d.prior_new = s->prior_new; d.lz_new = s->lz_new;
        d.wnew = s->wnew; d.where = s->where;
        d.par = s->par; d.root = s->root; d.alt = s->alt;
        d.lnew0 = s->out0; d.lnew1 = s->out1; d.gnode = gn; d.indicator = s->ind;
        d.h = s->h; d.loglik = s->ll; d.logpz = s->lz; d.prior = s->prior; d.accepted = s->sw_acc + (size_t)j * n;
        if (launch_gt_hdecide(d, stream) != hipSuccess) return engine_fail(e, PA_E_DEVICE, "height decision launch failed");
    }
    if (hipStreamSynchronize(stream) != hipSuccess) return engine_fail(e, PA_E_DEVICE, "height moves failed");
    if (accepted && hipMemcpy(accepted, s->sw_acc, kn * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return engine_fail(e, PA_E_DEVICE, "session download failed");
    if (kernel_ms) {
        double tot = 0;
        for (int j = 0; j < k; j++) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, s->ev[2 * j], s->ev[2 * j + 1]) != hipSuccess)
                return engine_fail(e, PA_E_DEVICE, "session events");
            tot += ms;
        }
        *kernel_ms = tot;
    }
    return gt_plan_stats(e, s->plan);
}

extern "C" pa_status pa_gt_session_read(pa_gt_session *s, double *gheight, double *loglik, double *logpz, double *prior)
{
    if (!s) return PA_E_INPUT;
    const size_t n = s->n, N = s->sp.N;
    hipError_t r = hipSuccess;
    if (gheight && r == hipSuccess) r = hipMemcpy(gheight, s->h, sizeof(double) * n * N, hipMemcpyDeviceToHost);
    if (loglik && r == hipSuccess) r = hipMemcpy(loglik, s->ll, sizeof(double) * 2 * n, hipMemcpyDeviceToHost);
    if (logpz && r == hipSuccess) r = hipMemcpy(logpz, s->lz, sizeof(double) * 2 * n, hipMemcpyDeviceToHost);
    if (prior && r == hipSuccess) r = hipMemcpy(prior, s->prior, sizeof(double) * n, hipMemcpyDeviceToHost);
    if (r != hipSuccess) return engine_fail(s->e, PA_E_DEVICE, "session download failed");
    return PA_OK;
}

extern "C" void pa_gt_session_destroy(pa_gt_session *s)
{
    if (!s) return;
    (void)hipDeviceSynchronize();
    delete s;
}
