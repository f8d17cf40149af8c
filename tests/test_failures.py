"""Per-element failure isolation and the chain failure bits.

The reference catches failures per element and marks an element failed when
its marginal likelihood is NaN (src/PhyloAcc-ST/main.cpp:458-499,
bpp_c2.cpp:339-348); it warns on rate proposals outside (1e-12, 1e8) and
skips them (bpp_c.cpp:1024-1041).  Here each chain carries PA_FLAG_* bits and
a status, and a batch with broken elements completes with every other
element's result unchanged.  Every condition is forced through inputs the
host controls, so every kernel runs to its normal end:

* zero-length sister branches with different bases at some site: P(t = 0)
  is the identity, the likelihood is 0 at any rates (-inf);
* a one-site element: the neutral-rate proposal is Gamma(r / s, s) with
  s = 10 / GG = 10, so about one proposal in twenty falls below 1e-12
  (out-of-range warning).  Adaptation is kept out of the run (it caps s at 1)
  and ROPT is 0 (ROPT 1 skips proposals below NLB before the range check);
* an element of gaps only: filtered when it is encoded;
* a NaN or infinite rate, or a NaN gene-branch length, in one element of a
  likelihood batch (a negative length acts as 0, as lengths below 1e-10 do
  in the reference); a model whose tree holds a non-finite length is
  refused.
"""
import re

import numpy as np
import pytest

from chain_checks import oracle_chains
from oracle import oracle_py as O
from phyloacc_amd import engine as E
from phyloacc_amd import synth

FAIL_BITS = E.PA_FLAG_NAN | E.PA_FLAG_POS_INF | E.PA_FLAG_SAMPLER


def _zero_cherry(w):
    """Set both branches of one cherry of the tree to length 0 and make the
    two species identical in every element but the first; returns the cherry's
    leaves."""
    t = w.tree
    cherry = next(v for v in range(t.S, t.N - 1) if t.children[v][0] < t.S and t.children[v][1] < t.S
                  and t.children[v][0] not in w.outgroup and t.children[v][1] not in w.outgroup)
    a, b = (int(x) for x in t.children[cherry])
    nw = t.newick
    for leaf in (a, b):
        nw = re.sub(rf"(?<![A-Za-z0-9_]){re.escape(t.names[leaf])}:[0-9.eE+-]+", f"{t.names[leaf]}:0", nw)
        t.dist[leaf] = 0.0
    t.newick = nw
    w.mod_text = synth.mod_text(w.pi, w.Q, t)
    al = w.align.copy()
    keep = slice(w.starts[0], w.starts[0] + w.sizes[0])
    first_b = al[b, keep].copy()
    al[b] = al[a]
    al[b, keep] = first_b                        # the first element keeps its differences
    assert np.any(al[a, keep] != al[b, keep])
    w.align = al
    return a, b


def _workload():
    """Config-2 elements plus, appended: a one-site element and a gaps-only
    element; the first element has the conflicting cherry."""
    w = synth.make_workload(2, seed=44, n_elements=6)
    a, b = _zero_cherry(w)
    S, G = w.align.shape
    one = np.ascontiguousarray(w.align[:, :1])
    gaps = np.full((S, 60), ord("-"), np.uint8)
    w.align = np.ascontiguousarray(np.concatenate([w.align, one, gaps], 1))
    w.starts = np.concatenate([w.starts, [G, G + 1]]).astype(np.int64)
    w.sizes = np.concatenate([w.sizes, [1, 60]]).astype(np.int64)
    return w


def test_oracle_flags_for_the_forced_conditions():
    kw = dict(burnin=10, mcmc=200, adapt_freq=1000, min_len=1, ropt=0)
    w = _workload()
    m = O.model_from_workload(w)
    oo = O.default_options(**kw)
    conflict = O.Element(m, oo, w.align, w.starts[0], w.sizes[0]).run(oo)
    for r in conflict:
        assert r["flags"] & 4 and not np.isfinite(r["max_loglik"])          # -inf at every state
    one = O.Element(m, oo, w.align, w.starts[6], 1).run(oo)
    assert any(r["flags"] & 16 for r in one)                                # out-of-range proposals
    assert all(np.isfinite(r["max_loglik"]) and not (r["flags"] & 7) for r in one)
    assert O.Element(m, O.default_options(**dict(kw, min_len=50)), w.align, w.starts[7], 60).filtered


def test_non_finite_tree_length_is_refused():
    w = synth.make_workload(2, seed=45, n_elements=2)
    w.mod_text = re.sub(r":[0-9.eE+-]+", ":nan", w.mod_text, count=1)
    with pytest.raises(E.PaError):
        E.Engine.from_workload(w)


@pytest.mark.gpu
def test_broken_elements_leave_the_batch_intact(require_gpu):
    kw = dict(burnin=10, mcmc=200, adapt_freq=1000, min_len=1, ropt=0)
    w = _workload()
    eng = E.Engine.from_workload(w, E.default_options(**kw))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    assert st[7] == E.PA_ELEM_FILTERED                                      # gaps only
    run = ids[st == E.PA_ELEM_OK]
    res = eng.run_chains(run, precision=E.PA_FP64, traces=True)
    m = O.model_from_workload(w)
    oo = O.default_options(**kw)
    refs = oracle_chains(m, oo, w, np.flatnonzero(st == E.PA_ELEM_OK))
    for k, i in enumerate(run):
        for ms in range(3):
            flags, status = int(res["flags"][k, ms]), int(res["status"][k, ms])
            assert flags == refs[k][ms]["flags"], (i, ms, flags, refs[k][ms]["flags"])
            if i == ids[0]:
                assert flags & E.PA_FLAG_NEG_INF and status == E.PA_ELEM_NUMERIC_FAIL
                continue
            assert status == E.PA_ELEM_OK and not (flags & FAIL_BITS)
            np.testing.assert_array_equal(res["Z"][k, ms], refs[k][ms]["Z"])
            np.testing.assert_allclose(res["loglik"][k, ms], refs[k][ms]["loglik"], rtol=1e-9)
    assert any(res["flags"][list(run).index(ids[6])] & E.PA_FLAG_PROPOSAL_RANGE)
    # the good elements alone give bit-identical results
    good = np.array([i for i in run if i != ids[0] and i != ids[6]], np.int32)
    alone = eng.run_chains(good, precision=E.PA_FP64, traces=True)
    sel = [list(run).index(i) for i in good]
    for f in ("max_loglik", "max_m", "final_nrate", "final_crate", "Z", "loglik"):
        np.testing.assert_array_equal(res[f][sel], alone[f])
    # FP32 chains flag the same conditions
    r32 = eng.run_chains(run, precision=E.PA_FP32)
    assert r32["status"][0].tolist() == [E.PA_ELEM_NUMERIC_FAIL] * 3
    assert np.all(r32["status"][1:] == E.PA_ELEM_OK)


@pytest.mark.gpu
def test_non_finite_inputs_stay_in_their_element(require_gpu):
    w = synth.make_workload(2, seed=45, n_elements=8)
    eng = E.Engine.from_workload(w)
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    ids = ids[st == E.PA_ELEM_OK]
    rng = np.random.default_rng(3)
    Z = np.full((len(ids), eng.N), -1, np.int32)
    for k, i in enumerate(ids):
        d = eng.element(int(i))
        Z[k, d["nodes"]] = 1
        Z[k, d["root"]] = 0
    rates = np.stack([rng.uniform(0.2, 0.6, len(ids)), rng.uniform(1.5, 3.0, len(ids))], 1)
    clean = {p: eng.loglik(ids, Z, rates, p) for p in (E.PA_FP64, E.PA_FP32)}
    bad = rates.copy()
    bad[1, 0] = np.nan
    bad[4, 0] = np.inf
    for p in (E.PA_FP64, E.PA_FP32):
        got = eng.loglik(ids, Z, bad, p)
        assert np.isnan(got[1]) and not np.isfinite(got[4])
        ok = [k for k in range(len(ids)) if k not in (1, 4)]
        np.testing.assert_array_equal(got[ok], clean[p][ok])
    # a NaN or negative gene-branch length in one gene-tree element
    gw = synth.make_gt_workload(seed=7, n_elements=10, S=16)
    geng = E.Engine.from_workload(gw, E.default_options(min_len=1))
    gids, gst = geng.add_elements_with_trees(gw.align, gw.starts, gw.sizes, gw.children())
    L = gw.lengths()
    ref = geng.loglik_lengths(gids, L, E.PA_FP64)
    Lb = L.copy()
    Lb[2, 0] = np.nan
    got = geng.loglik_lengths(gids, Lb, E.PA_FP64)
    assert np.isnan(got[2])
    np.testing.assert_array_equal(np.delete(got, 2), np.delete(ref, 2))
    Ln, L0 = L.copy(), L.copy()
    Ln[3, 0], L0[3, 0] = -0.5, 0.0
    np.testing.assert_array_equal(geng.loglik_lengths(gids, Ln, E.PA_FP64), geng.loglik_lengths(gids, L0, E.PA_FP64))
