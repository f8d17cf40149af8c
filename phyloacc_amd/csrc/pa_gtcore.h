// pa_gtcore.h -- a PhyloAcc-GT gene tree inside the species tree, as plain
// arrays, for host threads and device lanes alike: the embedding (the
// species of every gene node and where each gene lineage ends), the
// substitution length of every gene branch, the multispecies-coalescent prior
// (GTree::priorTree, src/PhyloAcc-GT/genetree.cpp:560-584), the missing-gene
// term (get_logpZ1, :1156-1190) and the window of the node-height move
// (Sample_BranchLen, :1658-1774).  Every function works on caller storage of
// N entries per array, so one code runs in pa_gtree.cpp's host threads and in
// the device session's one-lane-per-element kernels.
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define GT_HD __host__ __device__ inline
#else
#define GT_HD inline
#endif

namespace pa {
namespace gt {

// The species tree: parents (N above the root N - 1), heights, population
// sizes, post-order and children of the internal species.
struct SpView {
    int S, N;
    const int32_t *up;
    const double *H, *theta;
    const int32_t *post, *kid0, *kid1;
};

GT_HD double sp_top(const SpView &s, int x) { return s.up[x] < s.N ? s.H[s.up[x]] : INFINITY; }

// One gene tree: children [2N] (-1 at leaves; leaf i is species i), parents
// (N at the root), heights, a post-order of all N nodes, and its embedding:
// where[v], the species holding gene node v; end[v], the species where
// lineage v ends (its parent's; the species root for the gene root).
struct TreeView {
    int S, N, root;
    const int32_t *ch, *par, *order;
    const double *h;
    int32_t *where, *end;
};

// A gene node lies in the lowest species, on the climb from its first leaf,
// whose parent species is not lower than the node (GTree::findSp,
// genetree.cpp:2521-2534): its first child's species, climbed further.
GT_HD void embed(const SpView &s, TreeView &g)
{
    for (int q = 0; q < g.N; q++) {
        const int v = g.order[q];
        if (v < g.S) { g.where[v] = v; continue; }
        int x = g.where[g.ch[2 * v]];
        while (s.up[x] < s.N && s.H[s.up[x]] < g.h[v]) x = s.up[x];
        g.where[v] = x;
    }
    for (int v = 0; v < g.N; v++) g.end[v] = v == g.root ? s.N - 1 : g.where[g.par[v]];
}

// gene node v's height changed in place (between its children and parent):
// its species and the end of its children's lineages follow
GT_HD void repoint(const SpView &s, TreeView &g, int v)
{
    int x = g.where[g.ch[2 * v]];
    while (s.up[x] < s.N && s.H[s.up[x]] < g.h[v]) x = s.up[x];
    g.where[v] = x;
    g.end[g.ch[2 * v]] = g.end[g.ch[2 * v + 1]] = x;
}

// lineage v crosses species x (from where[v] up to end[v])
GT_HD bool crosses(const SpView &s, const TreeView &g, int v, int x)
{
    for (int y = g.where[v];; y = s.up[y]) {
        if (y == x) return true;
        if (y == g.end[v] || s.up[y] >= s.N) return false;
    }
}

// Substitution length of gene branch v: its pieces in the species it
// crosses, each at that species' rate (Z 2: r2, Z 1: r1, else 1); pieces
// shorter than 1e-8 are the identity of getlogTMc and add nothing.  The
// root's is 0.  Needs the embedding of every node, nothing of other branches.
GT_HD double branch_length(const SpView &s, const TreeView &g, int v, const int32_t *Z, double r1, double r2)
{
    if (v == g.root) return 0;
    const double hp = g.h[g.par[v]];
    double acc = 0;
    for (int x = g.where[v];; x = s.up[x]) {
        const bool last = x == g.end[v] || s.up[x] >= s.N;
        const double lo = fmax(g.h[v], s.H[x]), hi = last ? hp : sp_top(s, x);
        if (hi - lo >= 1e-8) acc += (hi - lo) * (Z[x] == 2 ? r2 : (Z[x] == 1 ? r1 : 1.0));
        if (last) break;
    }
    return acc;
}

GT_HD void lengths(const SpView &s, const TreeView &g, const int32_t *Z, double r1, double r2, double *len)
{
    for (int v = 0; v < g.N; v++) len[v] = branch_length(s, g, v, Z, r1, r2);
}

// One coalescence among k lineages after a wait hi - lo, and the wait of k
// lineages that do not coalesce before the top hi of the species; rate is
// 1 / theta and ldens log(2 rate).
GT_HD double coal_term(double ldens, double rate, double hi, double lo, int k)
{
    return ldens - (hi - lo) * k * (k - 1) * rate;
}
GT_HD double coal_tail(double rate, double hi, double lo, int k) { return (hi - lo) * k * (k - 1) * rate; }

// The coalescent prior: per internal species, the lineages entering from
// below (the sum over its child species of those leaving their tops; a leaf
// species sends its gene) coalesce at the gene nodes born there in height
// order.  Workspace: cnt [N], off [N + 1], leave [N] ints, hs [N] doubles.
GT_HD double prior(const SpView &s, const TreeView &g, int32_t *cnt, int32_t *off, int32_t *leave, double *hs)
{
    const int N = s.N;
    for (int x = 0; x < N; x++) cnt[x] = 0;
    for (int v = g.S; v < g.N; v++) cnt[g.where[v]]++;
    off[0] = 0;
    for (int x = 0; x < N; x++) off[x + 1] = off[x] + cnt[x];
    for (int x = 0; x < N; x++) cnt[x] = off[x];
    for (int v = g.S; v < g.N; v++) hs[cnt[g.where[v]]++] = g.h[v];
    double lp = 0;
    for (int q = 0; q < N; q++) {
        const int x = s.post[q];
        if (x < s.S) { leave[x] = 1; continue; }
        int k = leave[s.kid0[x]] + leave[s.kid1[x]];
        const int a = off[x], b = off[x + 1];
        leave[x] = k - (b - a);
        if (k == 0) continue;
        for (int i = a + 1; i < b; i++)                  // a few nodes per species: insertion sort
            for (int j = i; j > a && hs[j - 1] > hs[j]; j--) {
                const double t = hs[j]; hs[j] = hs[j - 1]; hs[j - 1] = t;
            }
        const double rate = 1.0 / s.theta[x], ldens = log(2.0 * rate);
        double t = s.H[x];
        for (int i = a; i < b; i++) {
            lp += coal_term(ldens, rate, hs[i], t, k);
            t = hs[i];
            k--;
        }
        if (k > 1 && s.up[x] != s.N) lp -= coal_tail(rate, s.H[s.up[x]], t, k);
    }
    return lp;
}

// Missing genes: the leaves' flags, an internal node when both children are.
GT_HD void missing(const TreeView &g, const uint8_t *leaf_missing, uint8_t *miss)
{
    for (int q = 0; q < g.N; q++) {
        const int v = g.order[q];
        miss[v] = v < g.S ? leaf_missing[v] : (uint8_t)(miss[g.ch[2 * v]] && miss[g.ch[2 * v + 1]]);
    }
}

// log P(missing | Z) over the leaves and the gene nodes born in internal
// species; P(missing | Z = 1) = c2m, P(missing | Z != 1) = n2m.  The four
// logarithms: log(1 - c2m), log(c2m), log(1 - n2m), log(n2m).
GT_HD void logpz_consts(double c2m, double n2m, double *lc)
{
    lc[0] = log(1 - c2m);
    lc[1] = log(c2m);
    lc[2] = log(1 - n2m);
    lc[3] = log(n2m);
}

// gene node v's term (false: v adds nothing)
GT_HD bool logpz_term(const SpView &s, const TreeView &g, const uint8_t *miss, const int32_t *Z, const double *lc,
                      int v, double *term)
{
    int x = v;
    if (v >= g.S) {
        x = g.where[v];
        if (x < s.S || g.h[v] < s.H[x]) return false;
    }
    *term = lc[(Z[x] == 1 ? 0 : 2) + (miss[v] ? 1 : 0)];
    return true;
}

GT_HD double logpz(const SpView &s, const TreeView &g, const uint8_t *miss, const int32_t *Z, double c2m, double n2m)
{
    double lc[4];
    logpz_consts(c2m, n2m, lc);
    double lp = 0, t;
    for (int q = 0; q < g.N; q++)
        if (logpz_term(s, g, miss, Z, lc, g.order[q], &t)) lp += t;
    return lp;
}

// The window of GTree::Sample_BranchLen for gene node v (genetree.cpp:
// 1658-1774): cnode is the child with the shorter branch (the second child on
// a tie); the node may go down to the bottom of css, the lowest species that
// both children's lineages cross, or to cnode, and up to its parent gene node
// when that is born in the same species, else to the top of its species; the
// gene root goes up to 10 theta_root / 2 (h(root) + 5 theta_root / 2 when the
// root is already higher).  Half-width delta * theta / 2 of v's species.
// status: 0 a proposal, 1 an empty window, -1 no species holds both lineages.
struct Window {
    int status;
    double lower1, upper1, half, lower, upper;
};

GT_HD Window window(const SpView &s, const TreeView &g, int v, double delta)
{
    Window w = {0, 0, 0, 0, 0, 0};
    const int a = g.ch[2 * v], b = g.ch[2 * v + 1];
    const bool first = g.h[v] - g.h[a] < g.h[v] - g.h[b];
    const int cnode = first ? a : b, sib = first ? b : a;
    const int ss = g.where[v];
    int css = -1;
    for (int x = g.where[cnode];; x = s.up[x]) {
        if (crosses(s, g, sib, x)) { css = x; break; }
        if (x == ss || s.up[x] >= s.N) break;
    }
    if (css < 0) { w.status = -1; return w; }
    const int p = g.par[v];
    w.upper1 = p < s.N && g.where[p] == ss ? g.h[p] : sp_top(s, ss);
    if (v == g.root) {
        w.upper1 = 10 * s.theta[s.N - 1] / 2;
        if (w.upper1 < g.h[v]) w.upper1 = g.h[v] + 5 * s.theta[s.N - 1] / 2;
    }
    w.lower1 = fmax(s.H[css], g.h[cnode]);
    w.half = delta * s.theta[ss] / 2;
    w.upper = fmin(w.upper1, g.h[v] + w.half);
    w.lower = fmax(w.lower1, g.h[v] - w.half);
    if (w.upper - w.lower == 0) w.status = 1;
    return w;
}

// The Metropolis-Hastings log ratio of an accepted-or-not height move once
// the proposal's prefix log-likelihoods lnew [2] are known.
GT_HD double height_ratio(const Window &w, double prop, int ind, const double *lnew, const double *loglik,
                          const double *logpz, double lz_new, double prior_new, double prior_cur)
{
    const double back = fmin(w.upper1, prop + w.half) - fmax(w.lower1, prop - w.half);
    double d = lnew[ind] - loglik[ind] + log(w.upper - w.lower) - log(back);
    if (logpz[ind] != 0) d += lz_new - logpz[ind];
    return d + prior_new - prior_cur;
}

}  // namespace gt
}  // namespace pa
