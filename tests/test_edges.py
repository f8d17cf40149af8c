"""Edge cases at the boundary: empty batches, ids the engine does not hold,
trees beyond the node limit, and (on the GPU) one batch mixing one-site
elements with the longest ones, against the oracle."""
import numpy as np
import pytest

from chain_checks import oracle_chains
from oracle import oracle_py as O
from phyloacc_amd import engine as E
from phyloacc_amd import synth


def _engine(n=4, seed=3, **kw):
    w = synth.make_workload(2, seed=seed, n_elements=n)
    eng = E.Engine.from_workload(w, E.default_options(**kw))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    return w, eng, ids[st == E.PA_ELEM_OK]


def test_empty_batches_do_nothing():
    """Every batched entry point takes n = 0 and returns at once (no device
    work, so this runs without a GPU)."""
    w, eng, ids = _engine()
    none = np.zeros(0, np.int32)
    r = eng.run_chains(none)
    assert r["max_loglik"].shape == (0, 3)
    assert eng.loglik(none, np.zeros((0, eng.N), np.int32), np.zeros((0, 2))).shape == (0,)
    assert eng.loglik_lengths(none, np.zeros((0, eng.N))).shape == (0,)
    assert eng.loglik_gbgc(none, np.zeros((0, eng.N)), np.zeros((0, eng.N)), np.zeros((0, 3))).shape == (0,)
    a, ll, acc = eng.sample_pi(none, np.zeros((0, eng.N)), 0, [], [])
    assert len(a) == len(ll) == len(acc) == 0
    ids2, st2 = eng.add_elements(w.align, np.zeros(0, np.int64), np.zeros(0, np.int32))
    assert len(ids2) == 0 and len(st2) == 0
    with pytest.raises(E.PaError) as ei:
        E.Batch(eng, none)
    assert ei.value.status == E.PA_E_INPUT


def test_unknown_or_filtered_ids_are_refused():
    w, eng, ids = _engine()
    for bad in ([-1], [len(ids) + 100]):
        with pytest.raises(E.PaError) as ei:
            eng.run_chains(np.array(bad, np.int32))
        assert ei.value.status == E.PA_E_INPUT
    gaps = np.full((eng.S, 30), ord("-"), np.uint8)
    fid, fst = eng.add_elements(gaps, np.array([0]), np.array([30]))
    assert fst[0] == E.PA_ELEM_FILTERED
    with pytest.raises(E.PaError) as ei:
        eng.loglik(fid, np.zeros((1, eng.N), np.int32), np.ones((1, 2)))
    assert ei.value.status == E.PA_E_INPUT and "filtered" in str(ei.value)


def test_tree_beyond_the_node_limit_is_refused():
    rng = np.random.default_rng(1)
    tree = synth.random_tree(300, 0.1, rng)          # 599 nodes > 512
    pi, Q = synth.gtr(rng)
    with pytest.raises(E.PaError):
        E.parse_mod(synth.mod_text(pi, Q, tree))


@pytest.mark.gpu
def test_one_site_and_longest_elements_in_one_batch(require_gpu):
    """Ragged extremes in one launch: one-site elements (MIN_LEN 1) beside the
    longest lengths of the heavy-tailed configuration, FP64 chains against the
    oracle's trajectories."""
    kw = dict(burnin=3, mcmc=7, adapt_freq=4, min_len=1)
    w = synth.make_workload(3, seed=3, lengths=[1, 1, 2, 9958, 9942])
    eng = E.Engine.from_workload(w, E.default_options(**kw))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    ok = st == E.PA_ELEM_OK
    assert ok.sum() >= 4
    ids, idx = ids[ok], np.flatnonzero(ok)
    refs = oracle_chains(O.model_from_workload(w), O.default_options(**kw), w, idx)
    r = eng.run_chains(ids, precision=E.PA_FP64, traces=True)
    for k in range(len(ids)):
        for ms in range(3):
            np.testing.assert_array_equal(r["Z"][k, ms], refs[k][ms]["Z"])
            np.testing.assert_allclose(r["loglik"][k, ms], refs[k][ms]["loglik"], rtol=1e-9)
            assert r["max_m"][k, ms] == refs[k][ms]["max_m"]
