"""GPU parity: the HIP engine against the CPU oracle, through the C ABI."""
import numpy as np
import pytest

from phyloacc_amd import engine as E
from phyloacc_amd import synth
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu


def _setup(config=1, n=6, seed=11, **kw):
    w = synth.make_workload(config, seed=seed, n_elements=n)
    eng = E.Engine.from_workload(w, E.default_options(**kw))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    m = O.model_from_workload(w)
    oo = O.default_options(**kw)
    return w, eng, ids, st, m, oo


def _random_Z(eng, ids, rng):
    Z = np.full((len(ids), eng.N), -1, np.int32)
    for k, i in enumerate(ids):
        d = eng.element(i)
        for s in d["nodes"]:
            Z[k, s] = rng.integers(0, 3)
        Z[k, d["root"]] = rng.integers(0, 2)
    return Z


def _setup_amb(config, n, seed, **kw):
    """As _setup, with IUPAC two-base codes and N sprinkled into the alignment."""
    w = _with_ambiguity(synth.make_workload(config, seed=seed, n_elements=n), seed + 100)
    eng = E.Engine.from_workload(w, E.default_options(**kw))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    return w, eng, ids, st, O.model_from_workload(w), O.default_options(**kw)


@pytest.mark.parametrize("config", [1, 2])
def test_loglik_fp64_and_fp32_vs_oracle(require_gpu, config):
    w, eng, ids, st, m, oo = _setup_amb(config, n=8, seed=5)
    assert any(eng.element(int(i))["codes"].max() >= 6 for i in ids)   # ambiguity codes reach the kernel
    rng = np.random.default_rng(0)
    Z = _random_Z(eng, ids, rng)
    rates = np.stack([rng.uniform(0.1, 0.9, len(ids)), rng.uniform(1.2, 4.0, len(ids))], 1)
    ref = np.array([O.Element(m, oo, w.align, w.starts[i], w.sizes[i]).loglik(Z[k], rates[k, 0], rates[k, 1])
                    for k, i in enumerate(ids)])
    l64 = eng.loglik(ids, Z, rates, E.PA_FP64)
    l32 = eng.loglik(ids, Z, rates, E.PA_FP32)
    print("fp64 rel err", np.max(np.abs(l64 - ref) / np.abs(ref)), "fp32 rel err", np.max(np.abs(l32 - ref) / np.abs(ref)))
    np.testing.assert_allclose(l64, ref, rtol=1e-9)     # FP64 path vs oracle: 1e-9 relative
    np.testing.assert_allclose(l32, ref, rtol=1e-6)     # FP32 rescaled partials: 1e-6 relative


def test_chain_fp64_matches_oracle_trajectory(require_gpu):
    kw = dict(burnin=20, mcmc=60, adapt_freq=25)
    w, eng, ids, st, m, oo = _setup(1, n=4, seed=7, **kw)
    res = eng.run_chains(ids, precision=E.PA_FP64, traces=True)
    for k, i in enumerate(ids):
        ref = O.Element(m, oo, w.align, w.starts[i], w.sizes[i]).run(oo)
        for ms in range(3):
            r = ref[ms]
            np.testing.assert_array_equal(res["Z"][k, ms], r["Z"])
            np.testing.assert_allclose(res["loglik"][k, ms], r["loglik"], rtol=1e-9)
            np.testing.assert_allclose(res["nrate"][k, ms], r["nrate"], rtol=1e-9)
            np.testing.assert_allclose(res["crate"][k, ms], r["crate"], rtol=1e-9)
            np.testing.assert_allclose(res["grate"][k, ms], r["grate"], rtol=1e-9)
            np.testing.assert_allclose(res["max_loglik"][k, ms], r["max_loglik"], rtol=1e-9)
            assert res["max_m"][k, ms] == r["max_m"]
            np.testing.assert_array_equal(res["max_Z"][k, ms], r["max_Z"])


def _with_missing(config, n, seed):
    """Workload whose elements have missing species: both leaves of a cherry
    (so their parent is missing too) and one more leaf keep only ~10% of
    their bases; the rest are gaps."""
    w = synth.make_workload(config, seed=seed, n_elements=n)
    t = w.tree
    cherry = next(v for v in range(t.S, t.N) if t.children[v][0] < t.S and t.children[v][1] < t.S)
    a, b = t.children[cherry]
    other = next(s for s in range(t.S) if s not in (a, b) and s not in w.outgroup)
    rng = np.random.default_rng(seed)
    al = w.align.copy()
    for s in (a, b, other):
        mask = rng.random(al.shape[1]) < 0.9
        al[s, mask] = ord("-")
    w.align = al
    return w


@pytest.mark.parametrize("config", [1, 2])
def test_missing_species_loglik_and_chain(require_gpu, config):
    w = _with_ambiguity(_with_missing(config, 6, 31), 32)
    kw = dict(burnin=15, mcmc=45, adapt_freq=20)
    eng = E.Engine.from_workload(w, E.default_options(**kw))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    ids = ids[st == E.PA_ELEM_OK]
    assert len(ids) > 0
    assert any(eng.element(int(i))["missing"].sum() >= 3 for i in ids)
    m = O.model_from_workload(w)
    oo = O.default_options(**kw)
    rng = np.random.default_rng(1)
    Z = _random_Z(eng, ids, rng)
    rates = np.stack([rng.uniform(0.1, 0.9, len(ids)), rng.uniform(1.2, 4.0, len(ids))], 1)
    ref = np.array([O.Element(m, oo, w.align, w.starts[i], w.sizes[i]).loglik(Z[k], rates[k, 0], rates[k, 1])
                    for k, i in enumerate(ids)])
    np.testing.assert_allclose(eng.loglik(ids, Z, rates, E.PA_FP64), ref, rtol=1e-9)
    np.testing.assert_allclose(eng.loglik(ids, Z, rates, E.PA_FP32), ref, rtol=1e-6)
    res = eng.run_chains(ids, precision=E.PA_FP64, traces=True)
    for k, i in enumerate(ids):
        r = O.Element(m, oo, w.align, w.starts[i], w.sizes[i]).run(oo)
        for ms in range(3):
            np.testing.assert_array_equal(res["Z"][k, ms], r[ms]["Z"])
            np.testing.assert_allclose(res["loglik"][k, ms], r[ms]["loglik"], rtol=1e-9)
            assert res["max_m"][k, ms] == r[ms]["max_m"]


@pytest.mark.skip(reason="size-class launches (64- and 256-thread workgroups on two streams) are switched off: "
                         "the chain kernel they ran on did not finish on the GPU and its cause is not found yet")
def test_chain_fp64_two_size_classes(require_gpu):
    """A batch mixing short and long elements runs as two concurrent size
    classes (64- and 256-thread workgroups); both match the oracle."""
    kw = dict(burnin=10, mcmc=30, adapt_freq=15)
    w, eng, ids, st, m, oo = _setup(2, n=12, seed=17, **kw)
    ids = ids[st == E.PA_ELEM_OK]
    P = np.array([eng.element(int(i))["P"] for i in ids])
    assert (P > 256).any() and (P <= 256).any()
    res = eng.run_chains(ids, precision=E.PA_FP64, traces=True)
    assert res["stats"]["n_launches"] == 2
    for k, i in enumerate(ids):
        ref = O.Element(m, oo, w.align, w.starts[i], w.sizes[i]).run(oo)
        for ms in range(3):
            np.testing.assert_array_equal(res["Z"][k, ms], ref[ms]["Z"])
            np.testing.assert_allclose(res["loglik"][k, ms], ref[ms]["loglik"], rtol=1e-9)
            assert res["max_m"][k, ms] == ref[ms]["max_m"]


def _with_ambiguity(w, seed):
    """Sprinkle IUPAC two-base codes and N into the alignment."""
    rng = np.random.default_rng(seed)
    al = w.align.copy()
    mask = rng.random(al.shape) < 0.03
    al[mask] = rng.choice(np.frombuffer(b"ryksmwn", np.uint8), size=int(mask.sum()))
    w.align = al
    return w


@pytest.mark.parametrize("variant", [
    dict(hyper_lrate2_a=0.0),                      # Dollo: no 1 -> 2 return after loss
    dict(ropt=2),
    dict(ropt=0),
    dict(r1_shared_uniform=0),
    dict(prune_tree=1, constomis=0.05),
    dict(trim_gap_percent=0.3, init_crate=0.3, init_nrate=2.0),
])
def test_chain_fp64_option_variants(require_gpu, variant):
    kw = dict(burnin=12, mcmc=36, adapt_freq=10, **variant)
    w = _with_missing(1, 5, 41)
    w = _with_ambiguity(w, 42)
    eng = E.Engine.from_workload(w, E.default_options(**kw))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    ids = ids[st == E.PA_ELEM_OK]
    assert len(ids) > 0
    m = O.model_from_workload(w)
    oo = O.default_options(**kw)
    res = eng.run_chains(ids, precision=E.PA_FP64, traces=True)
    for k, i in enumerate(ids):
        r = O.Element(m, oo, w.align, w.starts[i], w.sizes[i]).run(oo)
        for ms in range(3):
            np.testing.assert_array_equal(res["Z"][k, ms], r[ms]["Z"])
            np.testing.assert_allclose(res["loglik"][k, ms], r[ms]["loglik"], rtol=1e-9)
            np.testing.assert_allclose(res["nrate"][k, ms], r[ms]["nrate"], rtol=1e-9)
            np.testing.assert_allclose(res["crate"][k, ms], r[ms]["crate"], rtol=1e-9)
            np.testing.assert_allclose(res["l2rate"][k, ms], r[ms]["l2rate"], rtol=1e-9)


def test_loglik_scratch_regrows(require_gpu):
    """One engine: a small FP32 call, then an FP64 call over more and longer
    elements (the kept scratch grows, long elements get regions of their
    own), then the small call again; each against the oracle."""
    w, eng, ids, st, m, oo = _setup(2, n=24, seed=19)
    ids = ids[st == E.PA_ELEM_OK]
    idx = np.flatnonzero(st == E.PA_ELEM_OK)
    rng = np.random.default_rng(2)
    Z = _random_Z(eng, ids, rng)
    rates = np.stack([rng.uniform(0.1, 0.9, len(ids)), rng.uniform(1.2, 4.0, len(ids))], 1)
    ref = np.array([O.Element(m, oo, w.align, w.starts[i], w.sizes[i]).loglik(Z[k], rates[k, 0], rates[k, 1])
                    for k, i in enumerate(idx)])
    order = np.argsort([eng.element(int(i))["P"] for i in ids])
    small = order[:3]
    for sel, prec, tol in ((small, E.PA_FP32, 1e-6), (order, E.PA_FP64, 1e-9), (small, E.PA_FP32, 1e-6)):
        np.testing.assert_allclose(eng.loglik(ids[sel], Z[sel], rates[sel], prec), ref[sel], rtol=tol)
