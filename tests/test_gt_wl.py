"""Wang-Landau blocks of PhyloAcc-GT's Gibbs (src/PhyloAcc-GT/bpp_c.cpp:359-440):
the log-likelihood over two prefixes of each element's sites, batched, and the
indicator step that alternates between them while it learns their log
normalising constants.

The oracle's indicator step is pinned by what the method is for: started at
one pair of targets and then fed another, its constants' difference must
converge to the new targets' difference and the two indicators must be
visited equally (a sign error in the momentum or the renormalisation makes
them run away).  The prefix likelihoods are the oracle's gene-tree likelihood
over the prefix itself.  The engine's step is host code: it must give the
oracle's indicators and states exactly.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from phyloacc_amd import engine as E
from phyloacc_amd import synth
from test_gt_pi import _gt_engine, _oracle_args, SEED


@pytest.mark.parametrize("shift", [3.0, -5.0, 8.0])
def test_oracle_wl_learns_the_normalising_constants(shift):
    st = np.zeros(7)
    lz = np.array([0.0, -2.0])
    inds = []
    for m in range(4000):
        ll = np.array([-1000.0, -20.0]) if m == 0 else np.array([-1000.0 + shift, -20.0])
        inds.append(O.wl_step(SEED, m, ll, lz, st))
    want = (-1000.0 + shift) - (-20.0 - 2.0)
    assert abs((st[0] - st[1]) - want) < 0.05, (st, want)
    assert abs(np.mean(inds[-2000:]) - 0.5) < 0.05
    assert st[4] < 0.01                              # the learning rate has shrunk


def test_engine_wl_step_equals_oracle():
    w = synth.make_workload(1, seed=3, n_elements=2)
    eng = E.Engine.from_workload(w, E.default_options(seed=SEED))
    rng = np.random.default_rng(0)
    n = 5
    st = np.zeros((n, 7))
    ost = np.zeros((n, 7))
    base = rng.normal(-500, 50, size=(n, 2))
    for step in range(300):
        ll = base + (0 if step == 0 else rng.normal(0, 2, size=(n, 2)))
        lz = rng.normal(-3, 1, size=(n, 2))
        got = eng.wl_step(step, ll, lz, st)
        for i in range(n):
            row = ost[i].copy()
            want = O.wl_step(SEED, step, ll[i], lz[i], row)
            ost[i] = row
            assert got[i] == want, (step, i)
        np.testing.assert_array_equal(st, ost)


@pytest.mark.gpu
def test_prefix_likelihoods_vs_oracle(require_gpu):
    w, eng, ids, m, opts = _gt_engine(seed=14, n=10)
    L = w.lengths()
    rng = np.random.default_rng(3)
    GG = w.sizes.astype(int)
    lens0 = np.array([rng.integers(g // 2, g + 1) for g in GG], np.int32)
    lens1 = np.array([rng.integers(0, l0 + 1) for l0 in lens0], np.int32)
    lens0[0], lens1[0] = GG[0], GG[0]                          # full length: the plain call
    a = rng.uniform(0.1, 0.4, len(ids))
    for piA in (None, a):
        o0, o1 = eng.loglik_lengths_split(ids, L, lens0, lens1, piA=piA)
        f32 = eng.loglik_lengths_split(ids, L, lens0, lens1, piA=piA, precision=E.PA_FP32)
        for c in range(len(ids)):
            args = list(_oracle_args(w, c))
            for k, (lens, got, got32) in enumerate(((lens0, o0, f32[0]), (lens1, o1, f32[1]))):
                args[-1] = int(lens[c])
                if lens[c] == 0:
                    assert got[c] == 0.0 and got32[c] == 0.0
                    continue
                ref = (O.gt_loglik(m, *args, gap_prop=opts.gap_prop) if piA is None
                       else O.gt_loglik_pi(m, piA[c], *args, gap_prop=opts.gap_prop))
                assert abs(got[c] - ref) <= 1e-9 * abs(ref), (c, k, got[c], ref)
                assert abs(got32[c] - ref) <= 1e-6 * abs(ref), (c, k, got32[c], ref)
    full = eng.loglik_lengths(ids[:1], L[:1])
    o0, _ = eng.loglik_lengths_split(ids[:1], L[:1], lens0[:1], lens1[:1])
    np.testing.assert_allclose(o0, full, rtol=1e-12)
    with pytest.raises(E.PaError):
        eng.loglik_lengths_split(ids[:1], L[:1], [GG[0] + 1], [0])
