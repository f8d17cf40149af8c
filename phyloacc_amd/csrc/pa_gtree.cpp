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
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/phyloacc_gpu.h"
#include "pa_internal.h"
#include "pa_rng.h"

using namespace pa;

namespace {

struct Sp {
    int S, N;
    const int32_t *up;     // parent species, N above the root
    const double *H, *theta;
    double top(int s) const { return up[s] < N ? H[up[s]] : INFINITY; }
};

// One gene tree with its embedding.  where[v]: species branch holding gene
// node v; end[v]: the species where lineage v ends (its parent's, or the
// species root for the gene root).
struct Embed {
    const Sp *sp;
    int N, root = -1;
    const int32_t *ch, *par;
    const double *h;
    std::vector<int> where, end, order;

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
        std::vector<uint8_t> seen(N, 0);
        std::vector<std::pair<int, int>> st{{root, 0}};
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
            if (v < s.S) where[v] = v;
            else {
                int x = where[ch[2 * v]];
                while (s.up[x] < N && s.H[s.up[x]] < h[v]) x = s.up[x];
                where[v] = x;
            }
            order.push_back(v);
            st.pop_back();
        }
        if (leaves != s.S) return false;
        for (int v : order) end[v] = v == root ? N - 1 : where[par[v]];
        return true;
    }
    bool internal(int v) const { return v >= sp->S; }
    // lineage v crosses species x
    bool crosses(int v, int x) const
    {
        for (int s = where[v];; s = sp->up[s]) {
            if (s == x) return true;
            if (s == end[v] || sp->up[s] >= N) return false;
        }
    }
};

double prior_of(const Embed &g)
{
    const Sp &s = *g.sp;
    std::vector<int> enter(s.N, 0);
    std::vector<std::vector<double>> born(s.N);
    for (int v : g.order) {
        if (g.internal(v)) born[g.where[v]].push_back(g.h[v]);
        for (int x = g.where[v]; x != g.end[v] && s.up[x] < s.N;) {
            x = s.up[x];
            enter[x]++;
        }
    }
    double lp = 0;
    for (int x = s.S; x < s.N; x++) {
        int k = enter[x];
        if (k == 0) continue;
        std::sort(born[x].begin(), born[x].end());
        const double rate = 1.0 / s.theta[x], ldens = std::log(2.0 * rate);
        double t = s.H[x];
        for (double b : born[x]) {
            lp += ldens - (b - t) * k * (k - 1) * rate;
            t = b;
            k--;
        }
        if (k > 1 && s.up[x] != s.N) lp -= (s.H[s.up[x]] - t) * k * (k - 1) * rate;
    }
    return lp;
}

void missing_of(const Embed &g, const uint8_t *leaf_missing, std::vector<uint8_t> &miss)
{
    miss.assign(g.N, 0);
    for (int v : g.order) miss[v] = g.internal(v) ? (miss[g.ch[2 * v]] && miss[g.ch[2 * v + 1]]) : leaf_missing[v];
}

double logpz_of(const Embed &g, const std::vector<uint8_t> &miss, const int32_t *Z, double c2m, double n2m)
{
    const double lc[2] = {std::log(1 - c2m), std::log(c2m)}, ln[2] = {std::log(1 - n2m), std::log(n2m)};
    auto term = [&](int v, int x) { return Z[x] == 1 ? lc[miss[v]] : ln[miss[v]]; };
    double lp = 0;
    for (int v : g.order) {
        if (!g.internal(v)) lp += term(v, v);
        else if (g.where[v] >= g.sp->S && g.h[v] >= g.sp->H[g.where[v]]) lp += term(v, g.where[v]);
    }
    return lp;
}

// Substitution length of every gene branch: its pieces in the species it
// crosses, each at that species' rate (Z 2: r2, Z 1: r1, else 1); pieces
// shorter than 1e-8 are the identity of getlogTMc and add nothing.
void lengths_of(const Embed &g, const int32_t *Z, double r1, double r2, double *len)
{
    const Sp &s = *g.sp;
    for (int v = 0; v < g.N; v++) len[v] = 0;
    for (int v : g.order) {
        if (v == g.root) continue;
        const double hp = g.h[g.par[v]];
        double acc = 0;
        for (int x = g.where[v];; x = s.up[x]) {
            const bool last = x == g.end[v] || s.up[x] >= s.N;
            const double lo = std::max(g.h[v], s.H[x]), hi = last ? hp : s.top(x);
            if (hi - lo >= 1e-8) acc += (hi - lo) * (Z[x] == 2 ? r2 : (Z[x] == 1 ? r1 : 1.0));
            if (last) break;
        }
        len[v] = acc;
    }
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
    if (roots != 1) return engine_fail(e, PA_E_INPUT, "species tree must have one root");
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

// The window of one height move and where it came from.
struct HeightMove {
    int status = 0;          // 0: a proposal, 1: empty window, -1: no species holds both lineages
    double lower1 = 0, upper1 = 0, half = 0, lower = 0, upper = 0;
};

// GTree::Sample_BranchLen's window (src/PhyloAcc-GT/genetree.cpp:1658-1774):
// cnode is the child with the shorter branch (the second child on a tie);
// the node may go down to the bottom of css, the lowest species that both
// children's lineages cross, or to cnode, and up to its parent gene node when
// that is born in the same species, else to the top of its species; the
// gene root goes up to 10 theta_root / 2 (h(root) + 5 theta_root / 2 when the
// root is already higher).  The window's half-width is delta * theta / 2 of
// the node's species.
HeightMove height_window(const Embed &g, int v, double delta)
{
    const Sp &s = *g.sp;
    HeightMove w;
    const int a = g.ch[2 * v], b = g.ch[2 * v + 1];
    const bool first = g.h[v] - g.h[a] < g.h[v] - g.h[b];
    const int cnode = first ? a : b, sib = first ? b : a;
    const int ss = g.where[v];
    // lowest species on cnode's climb (up to ss) that sib's lineage also crosses
    int css = -1;
    for (int x = g.where[cnode];; x = s.up[x]) {
        if (g.crosses(sib, x)) { css = x; break; }
        if (x == ss || s.up[x] >= s.N) break;
    }
    if (css < 0) { w.status = -1; return w; }
    const int p = g.par[v];
    w.upper1 = p < s.N && g.where[p] == ss ? g.h[p] : s.top(ss);
    if (v == g.root) {
        w.upper1 = 10 * s.theta[s.N - 1] / 2;
        if (w.upper1 < g.h[v]) w.upper1 = g.h[v] + 5 * s.theta[s.N - 1] / 2;
    }
    w.lower1 = std::max(s.H[css], g.h[cnode]);
    w.half = delta * s.theta[ss] / 2;
    w.upper = std::min(w.upper1, g.h[v] + w.half);
    w.lower = std::max(w.lower1, g.h[v] - w.half);
    if (w.upper - w.lower == 0) w.status = 1;
    return w;
}

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
    std::vector<Embed> emb(n);
    std::vector<HeightMove> win(n);
    std::vector<std::vector<uint8_t>> miss(n);
    std::vector<double> prop(n, 0.0), rdraw(n, 0.0);
    const Key key{ev.key0, ev.key1};
    const uint32_t c3 = tagword(0, TAG_GT_BRANCH), c2 = (uint32_t)m;
    for (int i = 0; i < n; i++) {
        const int32_t *ch = gchildren + (size_t)i * 2 * N, *pa = gparent + (size_t)i * N;
        std::vector<uint8_t> lm(s.S);
        if ((st = element_gene_tree(e, ids[i], lm.data(), ch)) != PA_OK) return st;
        if (!emb[i].build(s, ch, pa, gheight + (size_t)i * N))
            return engine_fail(e, PA_E_INPUT, std::string(BAD_GENE_TREE) + " (element " + std::to_string(i) + ")");
        missing_of(emb[i], lm.data(), miss[i]);
        const int v = gnode[i], ind = indicator[i];
        if (v < s.S || v >= N || emb[i].where[v] < 0)
            return engine_fail(e, PA_E_INPUT, "gnode must be an internal gene node");
        if (ind != 0 && ind != 1) return engine_fail(e, PA_E_INPUT, "indicator must be 0 or 1");
        if (lens[2 * i + 1] < 0 || lens[2 * i + ind] < 0) return engine_fail(e, PA_E_INPUT, "negative prefix length");
        win[i] = height_window(emb[i], v, delta);
        if (win[i].status < 0)
            return engine_fail(e, PA_E_INPUT, "element " + std::to_string(i) + ": no species holds both children of gene node " +
                                                  std::to_string(v));
        if (win[i].status == 0) {
            const double u = slot_uniform(key, (uint32_t)slot, c2, c3, 0);
            prop[i] = std::max(u, 1e-5) * (win[i].upper - win[i].lower) + win[i].lower;
            rdraw[i] = slot_uniform(key, (uint32_t)slot, c2, c3, 1);
        }
        accepted[i] = 0;
    }
    // the proposals' lengths and their prefix likelihoods, one launch
    std::vector<int32_t> run, rid, l0, l1;
    std::vector<double> rlen, rpi;
    std::vector<std::vector<double>> newh(n);
    for (int i = 0; i < n; i++) {
        if (win[i].status != 0) continue;
        newh[i].assign(gheight + (size_t)i * N, gheight + (size_t)(i + 1) * N);
        newh[i][gnode[i]] = prop[i];
        Embed g2;
        g2.build(s, emb[i].ch, emb[i].par, newh[i].data());
        emb[i] = std::move(g2);
        run.push_back(i);
        rid.push_back(ids[i]);
        const int ind = indicator[i];
        l0.push_back(lens[2 * i + ind]);
        l1.push_back(std::min(lens[2 * i + 1], lens[2 * i + ind]));
        rlen.resize(rlen.size() + N);
        lengths_of(emb[i], Z + (size_t)i * N, rates[2 * i], rates[2 * i + 1], &rlen[rlen.size() - N]);
        if (piA) rpi.push_back(piA[i]);
    }
    const int k = (int)run.size();
    if (k == 0) return PA_OK;
    std::vector<double> o0(k), o1(k);
    if ((st = pa_loglik_lengths_split(e, rid.data(), k, rlen.data(), piA ? rpi.data() : nullptr, l0.data(), l1.data(),
                                      precision, o0.data(), o1.data())) != PA_OK)
        return st;
    for (int j = 0; j < k; j++) {
        const int i = run[j], ind = indicator[i];
        const HeightMove &w = win[i];
        const double lz = logpz_of(emb[i], miss[i], Z + (size_t)i * N, cons_to_mis, ncons_to_mis);
        const double pr = prior_of(emb[i]);
        const double back = std::min(w.upper1, prop[i] + w.half) - std::max(w.lower1, prop[i] - w.half);
        const double lnew[2] = {o0[j], o1[j]};
        double d = lnew[ind] - loglik[2 * i + ind] + std::log(w.upper - w.lower) - std::log(back);
        if (logpz[2 * i + ind] != 0) d += lz - logpz[2 * i + ind];
        d += pr - prior[i];
        if (std::log(rdraw[i]) < d) {
            accepted[i] = 1;
            gheight[(size_t)i * N + gnode[i]] = prop[i];
            loglik[2 * i] = lnew[0];
            loglik[2 * i + 1] = lnew[1];
            if (logpz[2 * i + 1] == 0) logpz[2 * i] = lz;
            else logpz[2 * i] = logpz[2 * i + 1] = lz;
            prior[i] = pr;
        }
    }
    return PA_OK;
}
