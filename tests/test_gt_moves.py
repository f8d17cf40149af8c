"""PhyloAcc-GT's gene tree inside the species tree, held as topology + heights:
the species of a gene node (GTree::findSp), the multispecies-coalescent prior
(GTree::priorTree, src/PhyloAcc-GT/genetree.cpp:560-584), the missing-gene
term (get_logpZ1, :1156-1190), the pieces a gene branch spends in species
branches, and the node-height move (Sample_BranchLen, :1658-1965).

The oracle is pinned by what the coalescent fixes:
  * the prior of one coalescence in a two-species tree is the exponential
    density in closed form;
  * over a three-species tree, exp(prior) integrates to 1 over every gene
    tree, embedding by embedding: 1 - exp(-2 T / theta) for the coalescence
    inside the internal branch and exp(-2 T / theta) / 3 for each of the three
    topologies formed in the root (the textbook concordance probability);
  * species assignments and pieces equal the ones recorded by an independent
    coalescent simulator while it built the tree;
  * the missing-gene term is a hand count;
  * with no data the height move must sample the prior: its chain is checked
    against the exponential laws of the coalescent where the move is
    reversible (all coalescences in the root species).
The engine's host functions must give the oracle's values.
"""
import numpy as np
import pytest
from scipy import stats

from oracle import oracle_py as O
from phyloacc_amd import synth

SEED = 11


def _three_species(theta3=0.15, theta4=0.25, tau1=0.05, tau2=0.1):
    # ((0, 1)3, 2)4
    parent = [3, 3, 4, 4, 5]
    height = [0, 0, 0, tau1, tau2]
    theta = [0, 0, 0, theta3, theta4]
    return O.sptree(3, parent, height, theta), tau1, tau2, theta3, theta4


def _gene3(pair, other, t1, t2):
    ch = [[-1, -1]] * 3 + [sorted(pair), sorted([3, other])]
    par = [0] * 5
    for a in pair:
        par[a] = 3
    par[3] = 4
    par[other] = 4
    par[4] = 5
    return O.GeneTree(ch, par, [0, 0, 0, t1, t2])


def test_prior_one_coalescence_closed_form():
    sp = O.sptree(2, [2, 2, 3], [0, 0, 0.3], [0, 0, 0.2])
    for t in (0.31, 0.5, 1.7):
        gt = O.GeneTree([[-1, -1], [-1, -1], [0, 1]], [2, 2, 3], [0, 0, t])
        assert O.gt_prior(sp, gt) == pytest.approx(np.log(2 / 0.2) - (t - 0.3) * 2 / 0.2, abs=1e-13)


def _gl(a, b, n=64):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (b - a) * x + 0.5 * (b + a), 0.5 * (b - a) * w


def test_prior_integrates_to_the_coalescent_probabilities():
    sp, tau1, tau2, th3, th4 = _three_species()
    X = 15 * max(th3, th4)
    # coalescence of 0 and 1 inside species 3, then with 2 in the root
    t1s, w1 = _gl(tau1, tau2)
    x2, w2 = _gl(0, X)
    mass_in = sum(wa * wb * np.exp(O.gt_prior(sp, _gene3((0, 1), 2, a, tau2 + b)))
                  for a, wa in zip(t1s, w1) for b, wb in zip(x2, w2))
    want_in = 1 - np.exp(-2 * (tau2 - tau1) / th3)
    assert mass_in == pytest.approx(want_in, rel=1e-9)
    # all three lineages reach the root: each of the three topologies
    x1, v1 = _gl(0, X)
    for pair, other in (((0, 1), 2), ((0, 2), 1), ((1, 2), 0)):
        mass = sum(wa * wb * np.exp(O.gt_prior(sp, _gene3(pair, other, tau2 + a, tau2 + a + b)))
                   for a, wa in zip(x1, v1) for b, wb in zip(x2, w2))
        assert mass == pytest.approx(np.exp(-2 * (tau2 - tau1) / th3) / 3, rel=1e-9), pair
    # concordance: 1 - 2/3 exp(-2T/theta)
    conc = mass_in + np.exp(-2 * (tau2 - tau1) / th3) / 3
    assert conc == pytest.approx(1 - 2 / 3 * np.exp(-2 * (tau2 - tau1) / th3), rel=1e-9)


def _sp_from_synth(tree, h, theta=2.0):
    N = tree.N
    par = np.where(tree.parent < 0, N, tree.parent)
    return O.sptree(tree.S, par, h, np.full(N, theta))


def _gene_from_synth(g, N):
    par = np.where(g.parent < 0, N, g.parent)
    return O.GeneTree(g.children, par, g.height)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_species_and_pieces_match_the_simulator(seed):
    rng = np.random.default_rng(seed)
    tree, h = synth.ultrametric_tree(9, 0.4, rng)
    sp = _sp_from_synth(tree, h)
    for c in range(20):
        g = synth.msc_gene_tree(tree, h, np.random.default_rng([seed, c]))
        gt = _gene_from_synth(g, tree.N)
        sg, ss, sl = O.gt_pieces(sp, gt)
        got = sorted(zip(sg.tolist(), ss.tolist(), sl.tolist()))
        want = sorted(zip(g.seg_gene.tolist(), g.seg_species.tolist(), g.seg_len.tolist()))
        assert [x[:2] for x in got] == [x[:2] for x in want]
        np.testing.assert_allclose([x[2] for x in got], [x[2] for x in want], rtol=0, atol=1e-12)
        # a gene node is born in the species of its children's last pieces
        for v in range(tree.S, tree.N):
            born = [s for (gg, s, _) in zip(g.seg_gene, g.seg_species, g.seg_len) if gg == g.children[v][0]][-1]
            assert O.gt_species(sp, gt, v) == born, v


def test_logpz_hand_count():
    sp, *_ = _three_species()
    c2m, n2m = 0.05, 0.3
    Z = [1, 0, 2, 1, 2]
    lc, ln_, lmc, lmn = np.log(1 - c2m), np.log(1 - n2m), np.log(c2m), np.log(n2m)
    gt = _gene3((0, 1), 2, 0.07, 0.3)             # node 3 born in species 3, node 4 in 4
    gt.miss[:] = [0, 0, 1, 0, 0]
    want = lc + ln_ + lmn + lc + ln_                # tips 0, 1, 2 (missing, Z=2); node 3 (Z3=1); node 4 (Z4=2)
    assert O.gt_logpz(sp, gt, Z, c2m, n2m) == pytest.approx(want, abs=1e-14)
    gt = _gene3((0, 1), 2, 0.2, 0.3)              # node 3 now born in the root species
    gt.miss[:] = [1, 1, 0, 1, 0]
    want = lmc + lmn + ln_ + lmn + ln_
    assert O.gt_logpz(sp, gt, Z, c2m, n2m) == pytest.approx(want, abs=1e-14)


def _no_data(S):
    rng = np.random.default_rng(0)
    model, _ = O.parse_mod(synth.mod_text(*synth.gtr(rng), synth.random_tree(S, 0.1, rng)))
    return model, np.full((S, 1), ord("a"), np.uint8)


def _run_heights(sp, gt, nodes, steps, delta=1.0):
    model, align = _no_data(sp.S)
    out = []
    ll = np.zeros(2)
    lz = np.zeros(2)
    prior = O.gt_prior(sp, gt)
    for m in range(steps):
        for b, v in enumerate(nodes):
            code, prior = O.gt_branch_len(model, sp, gt, np.ones(len(gt.par), np.int32), 0.5, 2.0, 0.01, 0.5, v,
                                          delta, 0, [0, 0], ll, lz, prior, SEED, m, b, align, 0, 1)
            assert code >= 0
        assert prior == pytest.approx(O.gt_prior(sp, gt), abs=1e-12)
        out.append(gt.h.copy())
    return np.array(out)


def test_height_move_samples_the_prior_one_coalescence():
    tau, theta = 0.3, 0.2
    sp = O.sptree(2, [2, 2, 3], [0, 0, tau], [0, 0, theta])
    gt = O.GeneTree([[-1, -1], [-1, -1], [0, 1]], [2, 2, 3], [0, 0, 0.45])
    H = _run_heights(sp, gt, [2], 30000)[2000::5, 2]
    lam, U = 2 / theta, 10 * theta / 2 - tau         # root window capped at 10 theta / 2
    cdf = lambda x: (1 - np.exp(-lam * (x - tau))) / (1 - np.exp(-lam * U))
    assert stats.kstest(H, cdf).pvalue > 1e-3
    assert H.min() > tau


def test_height_move_samples_the_prior_three_lineages_in_the_root():
    sp, tau1, tau2, th3, th4 = _three_species()
    gt = _gene3((0, 2), 1, 0.2, 0.35)              # (0, 2) can only meet in the root species
    H = _run_heights(sp, gt, [3, 4], 30000)[2000::5]
    t1, t2 = H[:, 3] - tau2, H[:, 4] - H[:, 3]
    assert stats.kstest(t1, "expon", args=(0, th4 / 6)).pvalue > 1e-3   # 3 lineages: rate 3*2/theta
    assert stats.kstest(t2, "expon", args=(0, th4 / 2)).pvalue > 1e-3   # 2 lineages: rate 2/theta
    assert abs(stats.pearsonr(t1, t2)[0]) < 0.05


def test_height_move_stays_inside_its_bounds():
    """Sample_BranchLen keeps the topology and the order of heights: every
    gene node stays above its children and below its parent."""
    rng = np.random.default_rng(4)
    tree, h = synth.ultrametric_tree(7, 0.4, rng)
    sp = _sp_from_synth(tree, h, theta=0.5)
    g = synth.msc_gene_tree(tree, h * 0.25, np.random.default_rng(9))
    gt = _gene_from_synth(g, tree.N)
    sp = _sp_from_synth(tree, h * 0.25, theta=0.5)
    ch0 = gt.ch.copy()
    H = _run_heights(sp, gt, list(range(tree.S, tree.N)), 300)
    assert np.array_equal(gt.ch, ch0)
    for row in H:
        for v in range(tree.S, tree.N):
            a, b = gt.ch[2 * v], gt.ch[2 * v + 1]
            assert row[v] > row[a] and row[v] > row[b]


def _u(seed, slot, m, tag, t):
    return O.lib().or_uniform(seed, t >> 2, slot, m, tag << 16, t & 3)


def test_height_move_known_answer():
    """One move worked by hand from the generator's two uniforms: window
    h +- delta * theta / 2 clipped to [tau, 10 theta / 2], the proposal
    max(u, 1e-5) * width + lower, and the decision log(r) < prior ratio +
    log(width_now / width_back)."""
    tau, theta, delta = 0.3, 0.2, 1.4
    sp = O.sptree(2, [2, 2, 3], [0, 0, tau], [0, 0, theta])
    model, align = _no_data(2)
    for m in range(40):
        h0 = 0.32 + 0.01 * m
        gt = O.GeneTree([[-1, -1], [-1, -1], [0, 1]], [2, 2, 3], [0, 0, h0])
        half = delta * theta / 2
        lo, hi = max(tau, h0 - half), min(1.0, h0 + half)
        u, r = _u(SEED, 5, m, 8, 0), _u(SEED, 5, m, 8, 1)
        prop = max(u, 1e-5) * (hi - lo) + lo
        lo2, hi2 = max(tau, prop - half), min(1.0, prop + half)
        lam = 2 / theta
        d = -lam * (prop - h0) + np.log(hi - lo) - np.log(hi2 - lo2)
        p0 = O.gt_prior(sp, gt)
        code, p1 = O.gt_branch_len(model, sp, gt, np.ones(3, np.int32), 0.5, 2.0, 0.01, 0.5, 2, delta, 0, [0, 0],
                                   np.zeros(2), np.zeros(2), p0, SEED, m, 5, align, 0, 1)
        assert code == int(np.log(r) < d), m
        assert gt.h[2] == pytest.approx(prop if code else h0, abs=1e-15)


def _tree_chain(sp, gt, steps, exact, align=None, lens=(0, 0), Tg_fn=None, seed=SEED):
    S, N = sp.S, len(gt.par)
    model, no_align = _no_data(S)
    align = no_align if align is None else align
    Z = np.ones(N, np.int32)
    ll, lz = np.zeros(2), np.zeros(2)
    if lens[0] > 0:
        site = O.gt_site_logliks(model, sp, gt, Z, 0.5, 2.0, align, 0, align.shape[1])
        ll[:] = site[:lens[0]].sum(), site[:lens[1]].sum()
    acc = 0
    states = []
    for m in range(steps):
        b = [v for v in range(N) if v != gt.root][m % (N - 1)]
        Tg = np.zeros((N, lens[0]), np.int8) if Tg_fn is None else Tg_fn(model, gt, m)
        acc += O.gt_sample_tree(model, sp, gt, Z, 0.5, 2.0, 0.01, 0.5, b, 0, lens, Tg, ll, lz, seed, m, 0, align, 0,
                                align.shape[1], exact=exact)
        if lens[0] > 0:
            site = O.gt_site_logliks(model, sp, gt, Z, 0.5, 2.0, align, 0, align.shape[1])
            assert ll == pytest.approx([site[:lens[0]].sum(), site[:lens[1]].sum()], rel=1e-12)
        states.append((gt.ch.copy(), gt.par.copy(), gt.h.copy()))
    return acc / steps, states


def _three_summary(states, tau2):
    conc = inbr = 0
    ages = []
    for ch, par, h in states:
        root = int(np.flatnonzero(par == 5)[0])
        inner = 7 - root                               # the internal node that is not the root
        pair = tuple(sorted(ch[2 * inner:2 * inner + 2]))
        conc += pair == (0, 1)
        inbr += pair == (0, 1) and h[inner] < tau2
        ages.append(h[root] - max(h[inner], tau2))
    n = len(states)
    return conc / n, inbr / n, np.array(ages)


def test_tree_move_with_complete_hastings_samples_the_prior():
    """With no data and the sibling choice's probability in the Hastings
    ratio, the topology move alone must reproduce the coalescent: the
    concordance probability, the chance that 0 and 1 meet inside their
    ancestral branch, and the root's exponential age."""
    sp, tau1, tau2, th3, th4 = _three_species()
    gt = _gene3((0, 1), 2, 0.07, 0.3)
    acc, states = _tree_chain(sp, gt, 60000, exact=True)
    assert 0.1 < acc < 0.6
    conc, inbr, ages = _three_summary(states[1000:], tau2)
    T = tau2 - tau1
    assert conc == pytest.approx(1 - 2 / 3 * np.exp(-2 * T / th3), abs=0.02)
    assert inbr == pytest.approx(1 - np.exp(-2 * T / th3), abs=0.02)
    assert stats.kstest(ages[::20], "expon", args=(0, th4 / 2)).pvalue > 1e-3


def test_tree_move_as_the_reference_writes_it_leaves_out_the_sibling_choice():
    """The reference's Hastings ratio omits the 1/2 of choosing the new
    sibling among a candidate's children, so its chain favours regrafting
    above the gene root: it spends clearly less time in the concordant
    topology than the coalescent says (the restatement keeps that)."""
    sp, tau1, tau2, th3, th4 = _three_species()
    gt = _gene3((0, 1), 2, 0.07, 0.3)
    _, states = _tree_chain(sp, gt, 40000, exact=False)
    conc, inbr, _ = _three_summary(states[1000:], tau2)
    assert conc < 1 - 2 / 3 * np.exp(-2 * (tau2 - tau1) / th3) - 0.025


def test_tree_move_keeps_a_valid_embedding():
    """With data: every state is a binary gene tree over all leaves whose
    nodes sit above the speciation of the species holding all their leaves
    and above their children."""
    rng = np.random.default_rng(5)
    tree, h = synth.ultrametric_tree(6, 0.4, rng)
    sp = _sp_from_synth(tree, h * 0.2, theta=0.4)
    g = synth.msc_gene_tree(tree, h * 0.2, np.random.default_rng(2))
    gt = _gene_from_synth(g, tree.N)
    align = np.frombuffer(b"acgtacgtta" + b"acgtacgata" + b"acctacgtta" + b"tcgtacgtta" + b"acgaacgtta" + b"acgtacttta", np.uint8).reshape(6, 10).copy()
    _, states = _tree_chain(sp, gt, 400, exact=False, align=align, lens=(10, 5),
                            Tg_fn=lambda model, gt, m: O.gt_sample_history(model, sp, gt, np.ones(tree.N, np.int32),
                                                                           0.5, 2.0, SEED, m, 10, align, 0, 10))
    changed = 0
    for ch, par, hh in states:
        changed += not np.array_equal(ch, states[0][0])
        root = int(np.flatnonzero(par == tree.N)[0])
        seen, stack = [], [root]
        while stack:
            v = stack.pop()
            seen.append(v)
            if v >= tree.S:
                a, b = ch[2 * v], ch[2 * v + 1]
                assert par[a] == v and par[b] == v and hh[a] < hh[v] and hh[b] < hh[v]
                stack += [a, b]
        assert sorted(seen) == list(range(tree.N))
        leaves = {}
        for v in reversed(seen):
            leaves[v] = {v} if v < tree.S else leaves[ch[2 * v]] | leaves[ch[2 * v + 1]]
        for v in range(tree.S, tree.N):
            x = min(leaves[v])
            while not leaves[v] <= set(_clade(tree, x)):
                x = tree.parent[x]
            assert hh[v] > h[x] * 0.2 - 1e-12
    assert changed > 0


def _clade(tree, x):
    if x < tree.S:
        return [x]
    return _clade(tree, tree.children[x][0]) + _clade(tree, tree.children[x][1])


def test_candidate_scores_known_answer():
    """Branch = leaf 0 of the concordant three-species tree, one site: the
    candidates are the gene root (its sampled base 2) at distance
    (H3 - 0) r0 + (H4 - H3) r3 + (h(root) - H4) r4 and the root candidate at
    ... + (curN - h(root)) r4; their scores are log of row x of P(len) times
    the leaf's indicator, and the pi-average of that message."""
    import scipy.linalg
    sp, tau1, tau2, th3, th4 = _three_species()
    gt = _gene3((0, 1), 2, 0.07, 0.3)
    model, _ = _no_data(3)
    align = np.frombuffer(b"c" + b"a" + b"g", np.uint8).reshape(3, 1).copy()
    Z = np.array([2, 1, 0, 1, 2], np.int32)
    cr, nr = 0.4, 2.5
    rate = {0: 1.0, 1: cr, 2: nr}
    Tg = np.array([[1], [0], [2], [3], [2]], np.int8)     # leaves' codes; node 3 -> 3, root 4 -> 2
    got = []
    O.gt_sample_tree(model, sp, gt, Z, cr, nr, 0.01, 0.5, 0, 0, [1, 1], Tg, np.zeros(2), np.zeros(2), SEED, 0, 0,
                     align, 0, 1, scores=got)
    Q = np.array([model.Q[i] for i in range(16)]).reshape(4, 4)
    pi = np.array([model.pi[i] for i in range(4)])
    cur = tau1 * rate[Z[0]] + (tau2 - tau1) * rate[Z[3]]
    curN = 10 * th4 / 2
    l_root = cur + (0.3 - tau2) * rate[Z[4]]
    l_N = cur + (curN - 0.3) * rate[Z[4]]
    Lb = np.array([0, 1.0, 0, 0])                          # leaf 0 holds c
    msg = lambda l: scipy.linalg.expm(Q * l) @ Lb
    want = [(4, np.log(msg(l_root)[2])), (5, np.log(pi @ msg(l_N)))]
    assert [c for c, _ in got] == [c for c, _ in want]
    np.testing.assert_allclose([s for _, s in got], [s for _, s in want], rtol=1e-10)


def test_history_draws_follow_the_posterior():
    """Update_Tg's draws at one site of a four-leaf gene tree against the
    exact posterior of the internal bases by enumeration (chi-square over
    many sweeps' independent uniforms)."""
    import scipy.linalg
    rng = np.random.default_rng(3)
    tree, h = synth.ultrametric_tree(4, 0.3, rng)
    sp = _sp_from_synth(tree, h * 0.5, theta=0.6)
    g = synth.msc_gene_tree(tree, h * 0.5, np.random.default_rng(8))
    gt = _gene_from_synth(g, tree.N)
    model, _ = _no_data(4)
    align = np.frombuffer(b"acgg", np.uint8).reshape(4, 1).copy()
    Z = np.array([0, 1, 2, 1, 2, 1, 1], np.int32)
    cr, nr = 0.5, 3.0
    Q = np.array([model.Q[i] for i in range(16)]).reshape(4, 4)
    pi = np.array([model.pi[i] for i in range(4)])
    ln = np.zeros(tree.N)
    sg, ss, sl = O.gt_pieces(sp, gt)
    for a, s_, l in zip(sg, ss, sl):
        ln[a] += l * {0: 1.0, 1: cr, 2: nr}[int(Z[s_])]
    P = {v: scipy.linalg.expm(Q * ln[v]) for v in range(tree.N)}
    obs = {0: 0, 1: 1, 2: 2, 3: 2}
    internal = list(range(tree.S, tree.N))
    root = gt.root
    import itertools
    post = {}
    for assign in itertools.product(range(4), repeat=len(internal)):
        x = dict(obs)
        x.update(zip(internal, assign))
        p = pi[x[root]]
        for v in range(tree.N):
            if v != root:
                p *= P[v][x[gt.par[v]], x[v]]
        post[assign] = p
    tot = sum(post.values())
    counts = {}
    n = 6000
    for m in range(n):
        T = O.gt_sample_history(model, sp, gt, Z, cr, nr, SEED, m, 1, align, 0, 1)
        assert list(T[:4, 0]) == [0, 1, 2, 2]
        key = tuple(int(T[v, 0]) for v in internal)
        counts[key] = counts.get(key, 0) + 1
    keys = [k for k in post if post[k] / tot * n >= 5]
    f_obs = np.array([counts.get(k, 0) for k in keys], float)
    f_exp = np.array([post[k] / tot * n for k in keys])
    f_exp *= f_obs.sum() / f_exp.sum()
    assert stats.chisquare(f_obs, f_exp).pvalue > 1e-3
