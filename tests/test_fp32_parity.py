"""The benched path -- chains over FP32 rescaled partials with FP64 sums --
against the FP64 oracle, through the C ABI.

Both sides draw from the same counter-based streams, so each model chain of
the engine can be compared state by state with the oracle's chain of the same
element (src/PhyloAcc-ST/bpp_c.cpp:346-519).  FP32 partials change the
log-likelihood by about 1e-7 relative; the gates are the ones the benchmark's
precision claim rests on: the per-model maximum log-likelihood within 1e-6
relative of the oracle's, identical per-branch Z posterior modes, and, for
chains whose Z trajectory is identical, a log-likelihood trace within 1e-6.
The FP64 path must reproduce the oracle's trajectories exactly.
"""
import numpy as np
import pytest

from chain_checks import compare, oracle_chains
from oracle import oracle_py as O
from phyloacc_amd import engine as E
from phyloacc_amd import synth

pytestmark = pytest.mark.gpu

FP32_RTOL = 1e-6     # FP32 partials vs the FP64 oracle, relative
FP64_RTOL = 1e-9     # FP64 path vs the FP64 oracle, relative

# Elements of the 50,000-element heavy-tailed configuration (Pareto 1.5 on
# [50, 10000] bp, 1-20% of branches accelerated): nine at evenly spaced length
# quantiles, and the three longest lengths of its seeded 50,000-element draw.
CONFIG3_LENGTHS = [int(x) for x in synth.pareto_quantiles(9)] + [9958, 9942, 9493]


def _engine(w, kw):
    eng = E.Engine.from_workload(w, E.default_options(**kw))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    ok = st == E.PA_ELEM_OK
    return eng, ids[ok], np.flatnonzero(ok)


def _random_Z(eng, ids, rng):
    Z = np.full((len(ids), eng.N), -1, np.int32)
    for k, i in enumerate(ids):
        d = eng.element(int(i))
        Z[k, d["nodes"]] = rng.integers(0, 3, size=len(d["nodes"]))
        Z[k, d["root"]] = rng.integers(0, 2)
    return Z


@pytest.mark.parametrize("config", [1, 2])
def test_fp32_chains_vs_oracle(require_gpu, config):
    kw = dict(burnin=10, mcmc=30, adapt_freq=15)
    w = synth.make_workload(config, seed=13, n_elements=8)
    eng, ids, idx = _engine(w, kw)
    m = O.model_from_workload(w)
    oo = O.default_options(**kw)
    refs = oracle_chains(m, oo, w, idx)
    res = eng.run_chains(ids, precision=E.PA_FP32, traces=True)
    rep = compare(res, refs, kw["burnin"], FP32_RTOL)
    print(f"config {config} fp32 chains vs oracle:", rep)
    assert rep["all_modes"], rep
    assert rep["max_rel_err_max_loglik"] <= FP32_RTOL, rep
    assert rep["max_rel_err_loglik_trace"] <= FP32_RTOL, rep
    assert np.all(res["status"] == E.PA_ELEM_OK)


def test_config3_sample_vs_oracle(require_gpu):
    """Heavy-tailed lengths: elements up to ~10,000 bp, thousands of patterns
    per element (many pattern blocks per thread), some of them in scratch
    regions of their own.  FP64 chains reproduce the oracle's trajectories;
    FP32 chains keep its Z posterior modes and maxima within 1e-6; fixed-(Z, r)
    likelihoods at both precisions."""
    kw = dict(burnin=3, mcmc=9, adapt_freq=4)       # adaptation runs inside the 12 states
    w = synth.make_workload(3, seed=3, lengths=CONFIG3_LENGTHS)
    eng, ids, idx = _engine(w, kw)
    assert len(ids) == len(CONFIG3_LENGTHS)
    P = np.array([eng.element(int(i))["P"] for i in ids])
    assert P.max() > 4000, P                          # far more patterns than threads
    m = O.model_from_workload(w)
    oo = O.default_options(**kw)
    refs = oracle_chains(m, oo, w, idx)

    r64 = eng.run_chains(ids, precision=E.PA_FP64, traces=True)
    assert r64["stats"]["scratch_bytes"] > 0
    for k in range(len(ids)):
        for ms in range(3):
            r = refs[k][ms]
            np.testing.assert_array_equal(r64["Z"][k, ms], r["Z"])
            np.testing.assert_allclose(r64["loglik"][k, ms], r["loglik"], rtol=FP64_RTOL)
            np.testing.assert_allclose(r64["nrate"][k, ms], r["nrate"], rtol=FP64_RTOL)
            np.testing.assert_allclose(r64["crate"][k, ms], r["crate"], rtol=FP64_RTOL)
            np.testing.assert_allclose(r64["max_loglik"][k, ms], r["max_loglik"], rtol=FP64_RTOL)
            assert r64["max_m"][k, ms] == r["max_m"]

    r32 = eng.run_chains(ids, precision=E.PA_FP32, traces=True)
    rep = compare(r32, refs, kw["burnin"], FP32_RTOL)
    print("config 3 sample fp32 chains vs oracle:", rep)
    assert rep["all_modes"], rep
    assert rep["max_rel_err_max_loglik"] <= FP32_RTOL, rep
    assert rep["max_rel_err_loglik_trace"] <= FP32_RTOL, rep

    rng = np.random.default_rng(4)
    Z = _random_Z(eng, ids, rng)
    rates = np.stack([rng.uniform(0.1, 0.5, len(ids)), rng.uniform(1.5, 4.0, len(ids))], 1)
    ref = np.array([O.Element(m, oo, w.align, w.starts[i], w.sizes[i]).loglik(Z[k], rates[k, 0], rates[k, 1])
                    for k, i in enumerate(idx)])
    l64 = eng.loglik(ids, Z, rates, E.PA_FP64)
    l32 = eng.loglik(ids, Z, rates, E.PA_FP32)
    print("config 3 sample loglik rel err: fp64 %.2e fp32 %.2e"
          % (np.max(np.abs(l64 - ref) / np.abs(ref)), np.max(np.abs(l32 - ref) / np.abs(ref))))
    np.testing.assert_allclose(l64, ref, rtol=FP64_RTOL)
    np.testing.assert_allclose(l32, ref, rtol=FP32_RTOL)


def test_per_element_launches_match_batched(require_gpu):
    """The plain path (one single-workgroup launch per element, every element
    in the workgroup's slot) gives the batched launch's results bit for bit."""
    kw = dict(burnin=4, mcmc=8, adapt_freq=5)
    w = synth.make_workload(3, seed=3, lengths=[60, 400, 2500, 9958])
    eng, ids, idx = _engine(w, kw)
    a = eng.run_chains(ids, precision=E.PA_FP64, traces=True)
    b = eng.run_chains(ids, precision=E.PA_FP64, traces=True, mode=E.PA_RUN_PER_ELEMENT)
    assert b["stats"]["n_launches"] == len(ids)
    for f in ("max_loglik", "max_m", "final_nrate", "final_crate", "Z", "loglik"):
        np.testing.assert_array_equal(a[f], b[f])
