"""PhyloAcc-GT's history draws and topology move on the device against the
oracle (GTree::Update_Tg with tosample = true, Sample_tree2 with
Remove_branch / Graft; src/PhyloAcc-GT/genetree.cpp:1416-1532,2138-2518,
737-923), whose own pins are in test_gt_moves.py.

FP64, element by element over several sweeps: identical histories, the
same accepted moves and trees (topology, heights), log-likelihoods within
1e-9; the height moves run between the topology moves on the regrafted
trees (by the host-driven move, or by a device session's sweep), so the
element's stored tree must follow every accepted regraft on the device as
well.  The Wang-Landau indicator alternates, element 0 has a
missing species, and a second run uses per-element base frequencies."""
import numpy as np
import pytest

from oracle import oracle_py as O
from phyloacc_amd import engine as E
from test_gt_branch import _gt_workload, _state, _species, _leaf_missing

SEED = 31


def _setup(use_pi):
    w = _gt_workload(seed=6, n=5, S=7)
    opts = E.default_options(min_len=1, seed=SEED)
    eng = E.Engine.from_workload(w, opts)
    ids, st = eng.add_elements_with_trees(w.align, w.starts, w.sizes, w.children())
    assert (st == E.PA_ELEM_OK).all()
    return w, eng, ids


@pytest.mark.gpu
@pytest.mark.parametrize("use_pi,heights", [(False, "host"), (True, "host"), (True, "session")])
def test_history_and_tree_moves_match_the_oracle(require_gpu, use_pi, heights):
    """heights: the height moves by pa_gt_branch_len, or by a device
    session's sweep (pa_gt_session_sweep) read back before the next sweep's
    history draw and regraft."""
    w, eng, ids = _setup(use_pi)
    m_or = O.parse_mod(w.mod_text)[0]
    sp, osp = _species(w)
    S, N, n = w.tree.S, w.tree.N, w.C
    rng = np.random.default_rng(2)
    piA = rng.uniform(0.15, 0.35, n) if use_pi else None
    c2m, n2m = 0.05, 0.4
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
    rates = np.stack([w.r1, w.r2], 1)
    # short prefixes keep the likelihood flat enough for regrafts to happen
    lens = np.stack([np.minimum(w.sizes, 24), np.minimum(w.sizes, 24) // 2], 1).astype(np.int32)
    full, half = eng.loglik_lengths_split(ids, w.lengths(), lens[:, 0], lens[:, 1], piA=piA)
    ll = np.stack([full, half], 1)
    lz = np.array([[O.gt_logpz(osp, og[c], w.Z[c], c2m, n2m)] * 2 for c in range(n)])
    oll = ll.copy()
    olz = lz.copy()
    moved = accepted = 0
    for m in range(8):
        ind = np.full(n, m % 2, np.int32)
        L = lens[np.arange(n), ind]
        hist = eng.gt_sample_history(sp, ids, ch, par, H, w.Z, rates, L, m, piA=piA)
        for c in range(n):
            ho = O.gt_sample_history(m_or, osp, og[c], w.Z[c], w.r1[c], w.r2[c], SEED, m, L[c], w.align, w.starts[c],
                                     w.sizes[c], piA=0.0 if piA is None else piA[c])
            np.testing.assert_array_equal(hist[c, :, :L[c]], ho, err_msg=f"history m={m} c={c}")
        branch = np.array([[v for v in range(N) if par[c, v] != N][(m * 3 + c) % (N - 1)] for c in range(n)], np.int32)
        ch0 = ch.copy()
        # one variant reads the histories the device kept, the other the caller's copy
        acc = eng.gt_sample_tree(sp, ids, ch, par, H, w.Z, rates, branch, ind, lens, None if use_pi else hist, ll,
                                 lz, m, 0, cons_to_mis=c2m, ncons_to_mis=n2m, piA=piA)
        for c in range(n):
            code = O.gt_sample_tree(m_or, osp, og[c], w.Z[c], w.r1[c], w.r2[c], c2m, n2m, branch[c], int(ind[c]),
                                    lens[c], hist[c, :, :L[c]], oll[c], olz[c], SEED, m, 0, w.align, w.starts[c],
                                    w.sizes[c], piA=0.0 if piA is None else piA[c])
            assert code == acc[c], (m, c)
            np.testing.assert_array_equal(ch[c].reshape(-1), og[c].ch)
            np.testing.assert_array_equal(par[c], og[c].par)
            np.testing.assert_array_equal(H[c], og[c].h)
            np.testing.assert_allclose(ll[c], oll[c], rtol=1e-9)
            np.testing.assert_allclose(lz[c], olz[c], rtol=1e-12)
        accepted += acc.sum()
        moved += (ch != ch0).any(axis=(1, 2)).sum()
        # height moves on the (possibly regrafted) trees
        prior = np.array([O.gt_prior(osp, og[c]) for c in range(n)])
        oprior = prior.copy()
        if heights == "session":
            ses = E.GtSession(eng, sp, ids, ch, par, H, w.Z, rates, ll, lz, prior, piA=piA, cons_to_mis=c2m,
                              ncons_to_mis=n2m)
            swept, _ = ses.sweep(np.arange(S, N, dtype=np.int32), ind, lens, m, 1)
            H[:], ll[:], lz[:], prior[:] = ses.read()
            ses.close()
        for b, v in enumerate(range(S, N)):
            if heights == "session":
                a2 = swept[b]
            else:
                a2 = eng.gt_branch_len(sp, ids, ch, par, H, w.Z, rates, v, ind, lens, ll, lz, prior, m, 1 + b,
                                       cons_to_mis=c2m, ncons_to_mis=n2m, piA=piA)
            for c in range(n):
                code, oprior[c] = O.gt_branch_len(m_or, osp, og[c], w.Z[c], w.r1[c], w.r2[c], c2m, n2m, v, 1.0,
                                                  int(ind[c]), lens[c], oll[c], olz[c], oprior[c], SEED, m, 1 + b,
                                                  w.align, w.starts[c], w.sizes[c],
                                                  piA=0.0 if piA is None else piA[c])
                assert code == a2[c], (m, v, c)
                if heights == "host":
                    np.testing.assert_array_equal(H[c], og[c].h)
                    np.testing.assert_allclose(ll[c], oll[c], rtol=1e-9)
        for c in range(n):
            np.testing.assert_array_equal(H[c], og[c].h)
            np.testing.assert_allclose(ll[c], oll[c], rtol=1e-9)
            np.testing.assert_allclose(lz[c], olz[c], rtol=1e-12)
            assert prior[c] == pytest.approx(oprior[c], rel=1e-12, abs=1e-12)
    assert accepted > 0 and moved > 0
    # the stored trees followed: a plain likelihood call at the final trees
    Lf = np.stack([E.gt_lengths(sp, ch[c], par[c], H[c], w.Z[c], w.r1[c], w.r2[c]) for c in range(n)])
    f0, f1 = eng.loglik_lengths_split(ids, Lf, w.sizes, w.sizes // 2, piA=piA)
    for c in range(n):
        site = None
        if piA is None:
            site = O.gt_site_logliks(m_or, osp, og[c], w.Z[c], w.r1[c], w.r2[c], w.align, w.starts[c], w.sizes[c])
            assert f0[c] == pytest.approx(site.sum(), rel=1e-9)


@pytest.mark.gpu
def test_tree_move_input_checks(require_gpu):
    w, eng, ids = _setup(False)
    sp, _ = _species(w)
    n, N = w.C, w.tree.N
    ch = np.stack([_state(w, c)[0] for c in range(n)])
    par = np.stack([_state(w, c)[1] for c in range(n)])
    H = np.stack([_state(w, c)[2] for c in range(n)])
    rates = np.stack([w.r1, w.r2], 1)
    lens = np.stack([w.sizes, w.sizes], 1).astype(np.int32)
    hist = eng.gt_sample_history(sp, ids, ch, par, H, w.Z, rates, lens[:, 0], 0)
    ll, lz = np.zeros((n, 2)), np.zeros((n, 2))
    with pytest.raises(E.PaError):                       # the root is no branch
        eng.gt_sample_tree(sp, ids, ch, par, H, w.Z, rates, N - 1, 0, lens, hist, ll, lz, 0, 0)
    with pytest.raises(E.PaError) as ei:                 # no device history for other ids
        eng.gt_sample_tree(sp, ids[::-1].copy(), ch[::-1].copy(), par[::-1].copy(), H[::-1].copy(), w.Z[::-1],
                           rates[::-1], 0, 0, lens, None, ll, lz, 0, 0)
    assert ei.value.status == E.PA_E_STATE
    with pytest.raises(E.PaError):                       # one element twice
        two = np.array([ids[0], ids[0]], np.int32)
        eng.gt_sample_tree(sp, two, ch[[0, 0]].copy(), par[[0, 0]].copy(), H[[0, 0]].copy(), w.Z[[0, 0]], rates[[0, 0]],
                           0, 0, lens[[0, 0]], hist[[0, 0]], np.zeros((2, 2)), np.zeros((2, 2)), 0, 0)
    other = ch.copy()
    other[0, N - 1] = other[0, N - 1][::-1]              # same tree, children swapped: accepted
    eng.gt_sample_history(sp, ids, other, par, H, w.Z, rates, lens[:, 0], 0)
    bad = ch.copy()
    a, b = bad[0, N - 1, 0], bad[0, N - 2, 0]
    bad[0, N - 1, 0], bad[0, N - 2, 0] = b, a            # not the stored tree
    with pytest.raises(E.PaError):
        eng.gt_sample_history(sp, ids, bad, par, H, w.Z, rates, lens[:, 0], 0)
