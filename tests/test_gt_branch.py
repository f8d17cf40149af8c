"""The engine's gene-tree state functions and the batched node-height move
(GTree::priorTree, get_logpZ1, Update_Tg's composed lengths, Sample_BranchLen;
src/PhyloAcc-GT/genetree.cpp:560-584,1156-1190,1330-1426,1658-1965) against
the oracle, whose own pins are in test_gt_moves.py.

Host functions: the same prior, missing-gene term and lengths as the oracle
on coalescent gene trees, and the lengths equal the simulator's own piece
sums.  Device move (FP64 likelihoods): element by element the oracle's
decisions and heights over whole sweeps, with the Wang-Landau prefixes, a
missing species, per-element base frequencies and an indicator that
alternates; log-likelihoods within 1e-9.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from phyloacc_amd import engine as E
from phyloacc_amd import synth

SEED = 23


def _gt_workload(seed=4, n=6, S=8, missing=True):
    w = synth.make_gt_workload(seed=seed, n_elements=n, S=S)
    if missing:
        # species 0 is all gaps in element 0: a missing gene
        w.align[0, w.starts[0]:w.starts[0] + w.sizes[0]] = ord("-")
    return w


def _state(w, c):
    """Element c's gene tree in substitution units: theta = 2 * scale puts
    the coalescent rate k(k-1)/theta per unit of height at k(k-1)/2 per
    coalescent unit, the simulator's."""
    g = w.genes[c]
    N = w.tree.N
    par = np.where(g.parent < 0, N, g.parent).astype(np.int32)
    return g.children.astype(np.int32), par, g.height * w.scale


def _species(w):
    N = w.tree.N
    par = np.where(w.tree.parent < 0, N, w.tree.parent)
    theta = np.full(N, 2 * w.scale)
    return (E.SpeciesTree.make(w.tree.S, par, w.sp_height * w.scale, theta),
            O.sptree(w.tree.S, par, w.sp_height * w.scale, theta))


def _leaf_missing(w, c, gap_prop=0.8):
    seg = w.align[:, w.starts[c]:w.starts[c] + w.sizes[c]]
    return ((seg == ord("-")).sum(1) > gap_prop * w.sizes[c]).astype(np.int32)


def test_host_functions_equal_the_oracle():
    w = _gt_workload(n=10, S=9)
    sp, osp = _species(w)
    rng = np.random.default_rng(0)
    for c in range(w.C):
        ch, par, h = _state(w, c)
        Z = w.Z[c]
        lm = _leaf_missing(w, c)
        og = O.GeneTree(ch, par, h)
        og.miss[:w.tree.S] = lm
        for v in range(w.tree.S, w.tree.N):
            og.miss[v] = og.miss[og.ch[2 * v]] and og.miss[og.ch[2 * v + 1]]
        assert E.gt_prior(sp, ch, par, h) == pytest.approx(O.gt_prior(osp, og), rel=1e-13, abs=1e-12)
        c2m, n2m = rng.uniform(0.01, 0.2), rng.uniform(0.2, 0.9)
        assert E.gt_logpz(sp, ch, par, h, lm, Z, c2m, n2m) == pytest.approx(O.gt_logpz(osp, og, Z, c2m, n2m),
                                                                             rel=1e-13)
        got = E.gt_lengths(sp, ch, par, h, Z, w.r1[c], w.r2[c])
        np.testing.assert_allclose(got, w.lengths()[c], rtol=1e-12, atol=1e-15)
        sg, ss, sl = O.gt_pieces(osp, og)
        want = E.gene_lengths(w.tree.N, sg, ss, sl, Z, w.r1[c], w.r2[c])
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)


def test_host_functions_refuse_bad_trees():
    w = _gt_workload(n=1, S=5)
    sp, _ = _species(w)
    ch, par, h = _state(w, 0)
    bad = h.copy()
    bad[w.tree.N - 1] = -1.0                       # root below its children
    with pytest.raises(E.PaError):
        E.gt_prior(sp, ch, par, bad)
    par2 = par.copy()
    par2[0] = w.tree.N                             # two roots
    with pytest.raises(E.PaError):
        E.gt_lengths(sp, ch, par2, h, w.Z[0], 0.3, 2.0)


@pytest.mark.gpu
@pytest.mark.parametrize("use_pi", [False, True])
def test_branch_len_sweeps_match_the_oracle(require_gpu, use_pi):
    w = _gt_workload()
    opts = E.default_options(min_len=1, seed=SEED)
    eng = E.Engine.from_workload(w, opts)
    ids, st = eng.add_elements_with_trees(w.align, w.starts, w.sizes, w.children())
    assert (st == E.PA_ELEM_OK).all()
    m_or = O.parse_mod(w.mod_text)[0]
    sp, osp = _species(w)
    S, N, n = w.tree.S, w.tree.N, w.C
    rng = np.random.default_rng(1)
    piA = rng.uniform(0.15, 0.35, n) if use_pi else None
    c2m, n2m, delta = 0.05, 0.4, 1.0
    ch = np.zeros((n, N, 2), np.int32)
    par = np.zeros((n, N), np.int32)
    H = np.zeros((n, N))
    og = []
    for c in range(n):
        ch[c], par[c], H[c] = _state(w, c)
        g = O.GeneTree(ch[c], par[c], H[c])
        g.miss[:S] = _leaf_missing(w, c)
        for v in range(S, N):
            g.miss[v] = g.miss[g.ch[2 * v]] and g.miss[g.ch[2 * v + 1]]
        og.append(g)
    assert og[0].miss[0] == 1
    rates = np.stack([w.r1, w.r2], 1)
    lens = np.stack([w.sizes, w.sizes // 2], 1).astype(np.int32)
    ll = np.zeros((n, 2))
    lz = np.zeros((n, 2))
    prior = np.zeros(n)
    oll, olz, oprior = [], [], []
    for c in range(n):
        site = O.gt_site_logliks(m_or, osp, og[c], w.Z[c], w.r1[c], w.r2[c], w.align, w.starts[c], w.sizes[c]) \
            if not use_pi else None
        if use_pi:
            full, half = eng.loglik_lengths_split(ids[c:c + 1], w.lengths()[c:c + 1], lens[c:c + 1, 0],
                                                  lens[c:c + 1, 1], piA=piA[c:c + 1])
            ll[c] = full[0], half[0]
        else:
            ll[c] = site.sum(), site[:lens[c, 1]].sum()
        lz[c] = O.gt_logpz(osp, og[c], w.Z[c], c2m, n2m)
        prior[c] = O.gt_prior(osp, og[c])
        oll.append(ll[c].copy())
        olz.append(lz[c].copy())
        oprior.append(prior[c])
    total = 0
    for m in range(4):
        ind = np.full(n, m % 2, np.int32)
        for b, v in enumerate(range(S, N)):
            acc = eng.gt_branch_len(sp, ids, ch, par, H, w.Z, rates, v, ind, lens, ll, lz, prior, m, b,
                                    delta=delta, cons_to_mis=c2m, ncons_to_mis=n2m, piA=piA)
            for c in range(n):
                code, oprior[c] = O.gt_branch_len(m_or, osp, og[c], w.Z[c], w.r1[c], w.r2[c], c2m, n2m, v, delta,
                                                  int(ind[c]), lens[c], oll[c], olz[c], oprior[c], SEED, m, b,
                                                  w.align, w.starts[c], w.sizes[c],
                                                  piA=0.0 if piA is None else piA[c])
                assert code == acc[c], (m, v, c)
                np.testing.assert_array_equal(H[c], og[c].h)
                np.testing.assert_allclose(ll[c], oll[c], rtol=1e-9)
                np.testing.assert_allclose(lz[c], olz[c], rtol=1e-12)
                assert prior[c] == pytest.approx(oprior[c], rel=1e-12, abs=1e-12)
            total += acc.sum()
    assert 0 < total < 4 * n * (N - S)


@pytest.mark.gpu
@pytest.mark.parametrize("use_pi,S,precs", [(False, 9, "ddd"), (True, 9, "ddd"), (False, 41, "ddd"),
                                            (True, 9, "dff"), (True, 41, "fff")])
def test_device_session_equals_the_host_moves(require_gpu, use_pi, S, precs):
    """The device-resident height moves (pa_gt_session_*) give, move for move,
    the decisions, heights and state of pa_gt_branch_len, and read back the
    same final state, the prior and missing-gene term to the bit; both paths
    are held to the oracle above.  S = 41 puts more gene nodes than a wave
    has lanes, and paths deeper than the path pass's loads in flight (4 in
    FP64, 8 in FP32).  precs: each sweep's precision (d FP64, f FP32); a
    change rebuilds the session's resident partials."""
    w = _gt_workload(seed=8, n=7, S=S)
    opts = E.default_options(min_len=1, seed=SEED)
    eng = E.Engine.from_workload(w, opts)
    ids, st = eng.add_elements_with_trees(w.align, w.starts, w.sizes, w.children())
    assert (st == E.PA_ELEM_OK).all()
    sp, osp = _species(w)
    S, N, n = w.tree.S, w.tree.N, w.C
    rng = np.random.default_rng(4)
    piA = rng.uniform(0.15, 0.35, n) if use_pi else None
    ch = np.zeros((n, N, 2), np.int32)
    par = np.zeros((n, N), np.int32)
    H = np.zeros((n, N))
    for c in range(n):
        ch[c], par[c], H[c] = _state(w, c)
    rates = np.stack([w.r1, w.r2], 1)
    lens = np.stack([w.sizes, w.sizes // 3], 1).astype(np.int32)
    f0, f1 = eng.loglik_lengths_split(ids, w.lengths(), lens[:, 0], lens[:, 1], piA=piA)
    ll = np.stack([f0, f1], 1)
    lz = np.full((n, 2), -3.0)
    prior = np.array([E.gt_prior(sp, ch[c], par[c], H[c]) for c in range(n)])
    ses = E.GtSession(eng, sp, ids, ch, par, H, w.Z, rates, ll, lz, prior, piA=piA, cons_to_mis=0.05,
                      ncons_to_mis=0.4)
    total = 0
    for m in range(3):
        ind = np.full(n, m % 2, np.int32)
        prec = E.PA_FP64 if precs[m] == "d" else E.PA_FP32
        for b, v in enumerate(range(S, N)):
            a_host = eng.gt_branch_len(sp, ids, ch, par, H, w.Z, rates, v, ind, lens, ll, lz, prior, m, b,
                                       cons_to_mis=0.05, ncons_to_mis=0.4, piA=piA, precision=prec)
            a_dev = ses.branch_len(v, ind, lens, m, b, precision=prec)
            np.testing.assert_array_equal(a_dev, a_host, err_msg=f"m={m} v={v}")
            total += a_host.sum()
        h2, ll2, lz2, pr2 = ses.read()
        np.testing.assert_array_equal(h2, H)
        np.testing.assert_allclose(ll2, ll, rtol=1e-12)
        np.testing.assert_array_equal(lz2, lz)
        np.testing.assert_array_equal(pr2, prior)
    assert total > 0
    ses.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [E.PA_FP64, E.PA_FP32])
def test_session_sweep_equals_single_moves(require_gpu, prec):
    """pa_gt_session_sweep (every internal gene node in turn, queued at
    once) leaves two sessions from the same state in the same place as one
    pa_gt_session_branch_len call per node: the same decisions, heights and
    state, to the bit; a second sweep continues from there."""
    w = _gt_workload(seed=11, n=6, S=33)
    opts = E.default_options(min_len=1, seed=SEED)
    eng = E.Engine.from_workload(w, opts)
    ids, st = eng.add_elements_with_trees(w.align, w.starts, w.sizes, w.children())
    assert (st == E.PA_ELEM_OK).all()
    sp, _ = _species(w)
    S, N, n = w.tree.S, w.tree.N, w.C
    ch = np.zeros((n, N, 2), np.int32)
    par = np.zeros((n, N), np.int32)
    H = np.zeros((n, N))
    for c in range(n):
        ch[c], par[c], H[c] = _state(w, c)
    rates = np.stack([w.r1, w.r2], 1)
    lens = np.stack([w.sizes, w.sizes // 2], 1).astype(np.int32)
    piA = np.random.default_rng(5).uniform(0.2, 0.3, n)
    f0, f1 = eng.loglik_lengths_split(ids, w.lengths(), lens[:, 0], lens[:, 1], piA=piA, precision=prec)
    ll = np.stack([f0, f1], 1)
    lz = np.full((n, 2), -2.0)
    prior = np.array([E.gt_prior(sp, ch[c], par[c], H[c]) for c in range(n)])
    one = E.GtSession(eng, sp, ids, ch, par, H, w.Z, rates, ll, lz, prior, piA=piA)
    many = E.GtSession(eng, sp, ids, ch, par, H, w.Z, rates, ll, lz, prior, piA=piA)
    rng = np.random.default_rng(6)
    for m in range(2):
        ind = np.full(n, m % 2, np.int32)
        g = np.stack([rng.permutation(np.arange(S, N)) for _ in range(n)], 1).astype(np.int32)   # [k, n]
        acc1 = np.stack([one.branch_len(g[j], ind, lens, m, 3 + j, precision=prec) for j in range(len(g))])
        acc2, ms = many.sweep(g, ind, lens, m, 3, precision=prec)
        np.testing.assert_array_equal(acc2, acc1)
        assert ms > 0
        for a, b in zip(one.read(), many.read()):
            np.testing.assert_array_equal(b, a)
    assert acc1.sum() > 0
    one.close()
    many.close()
