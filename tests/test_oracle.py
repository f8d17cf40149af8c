"""Pins the CPU oracle before it is trusted as the parity checker.

Every check compares the oracle with something it does not share code
with: rocRAND / Random123 Philox vectors, scipy.linalg.expm, closed forms
(JC69/F81), brute-force summation over ancestral states, scipy.stats
distributions, and structural invariants of the likelihood.
"""
import json
import os

import numpy as np
import pytest
import scipy.stats

from oracle import oracle_py as O
from phyloacc_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "golden.npz"))


def test_philox_known_answers():
    # vectors from rocRAND's philox4x32_10_engine (tests/golden/philox_kat.cpp);
    # the first three are also Random123's published known-answer vectors
    kats = json.load(open(os.path.join(HERE, "golden", "philox_kat.json")))
    for v in kats:
        assert O.philox(v["ctr"], v["key"]) == v["out"]
    assert kats[0]["out"] == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]


def test_expm_against_scipy():
    for Q, t, P in zip(G["expm_Q"], G["expm_t"], G["expm_P"]):
        out = np.zeros(16)
        q = np.ascontiguousarray(Q.reshape(-1))
        O.lib().or_expm(q.ctypes.data_as(O.C.POINTER(O.C.c_double)), float(t), out.ctypes.data_as(O.C.POINTER(O.C.c_double)))
        np.testing.assert_allclose(out.reshape(4, 4), P, rtol=1e-11, atol=1e-14)


def _expm(Q, t):
    out = np.zeros(16)
    q = np.ascontiguousarray(np.asarray(Q, np.float64).reshape(-1))
    O.lib().or_expm(q.ctypes.data_as(O.C.POINTER(O.C.c_double)), float(t), out.ctypes.data_as(O.C.POINTER(O.C.c_double)))
    return out.reshape(4, 4)


def test_expm_closed_forms_and_invariants():
    # F81 / JC69: P_ii = pi_i + (1 - pi_i) e^{-beta t}, P_ij = pi_j (1 - e^{-beta t})
    pi = np.array([0.1, 0.2, 0.3, 0.4])
    beta = 1.0 / (1.0 - np.sum(pi ** 2))
    Q = beta * (np.ones((4, 1)) * pi[None, :])
    np.fill_diagonal(Q, 0)
    np.fill_diagonal(Q, -Q.sum(1))
    for t in (0.0, 0.05, 0.7, 3.0):
        P = _expm(Q, t)
        e = np.exp(-beta * t)
        ref = np.ones((4, 1)) * pi[None, :] * (1 - e) + np.eye(4) * e
        np.testing.assert_allclose(P, ref, rtol=1e-12, atol=1e-15)
    rng = np.random.default_rng(1)
    pi, Q = synth.gtr(rng)
    for t in (1e-6, 0.3, 5.0):
        P = _expm(Q, t)
        np.testing.assert_allclose(P.sum(1), 1.0, atol=1e-13)           # rows sum to 1
        np.testing.assert_allclose(pi @ P, pi, atol=1e-13)              # pi P = pi
        np.testing.assert_allclose(pi[:, None] * P, (pi[:, None] * P).T, atol=1e-13)  # reversibility
    np.testing.assert_allclose(_expm(Q, 0.0), np.eye(4), atol=0)       # P(0) = I


def _brute_case(i):
    mod = str(G[f"brute{i}_mod"])
    rows = [str(r) for r in G[f"brute{i}_rows"]]
    m, _ = O.parse_mod(mod)
    S = m.S
    O.lib().or_set_groups(O.C.byref(m), None, 0, None, 0)
    align = np.array([np.frombuffer(r.encode(), np.uint8) for r in rows])
    return m, align, G[f"brute{i}_Z"], G[f"brute{i}_rates"], float(G[f"brute{i}_loglik"])


@pytest.mark.parametrize("i", range(6))
def test_loglik_brute_force(i):
    m, align, Z, rates, ref = _brute_case(i)
    opts = O.default_options(min_len=1)
    e = O.Element(m, opts, align, 0, align.shape[1])
    assert not e.filtered
    ll = e.loglik(Z, rates[0], rates[1])
    assert abs(ll - ref) <= 1e-12 * abs(ref), (ll, ref)


def test_loglik_invariants():
    w = synth.small_workload(S=6, C=1, length=60, seed=4)
    m = O.model_from_workload(w)
    opts = O.default_options(min_len=1)
    Z = np.ones(m.N, np.int32)
    Z[m.N - 1] = 0
    Z[2] = 2
    base = O.Element(m, opts, w.align, 0, 60).loglik(Z, 0.4, 2.5)
    # site permutation invariance
    perm = np.random.default_rng(0).permutation(60)
    e2 = O.Element(m, opts, np.ascontiguousarray(w.align[:, perm]), 0, 60).loglik(Z, 0.4, 2.5)
    assert abs(e2 - base) < 1e-10 * abs(base)
    # duplicating every column doubles the log-likelihood (pattern weights)
    dup = np.ascontiguousarray(np.concatenate([w.align, w.align], 1))
    e3 = O.Element(m, opts, dup, 0, 120).loglik(Z, 0.4, 2.5)
    assert abs(e3 - 2 * base) < 1e-10 * abs(base)
    # an all-gap column adds exactly log 1
    gapc = np.full((m.S, 1), ord("-"), np.uint8)
    e4 = O.Element(m, opts, np.ascontiguousarray(np.concatenate([w.align, gapc], 1)), 0, 61).loglik(Z, 0.4, 2.5)
    assert abs(e4 - base) < 1e-12 * abs(base)
    # all rates -> 0: a constant column scores log pi_x, a variable one -inf
    const = np.full((m.S, 3), ord("c"), np.uint8)
    Z0 = np.ones(m.N, np.int32)
    Z0[m.N - 1] = 1
    tiny = O.Element(m, opts, const, 0, 3).loglik(Z0, 1e-14, 1e-14)
    assert abs(tiny - 3 * np.log(w.pi[1])) < 1e-9


def test_gamma_beta_samplers():
    seed = 77
    for a, b in ((0.5, 2.0), (3.0, 0.2), (40.0, 0.05)):
        xs = np.array([O.lib().or_gamma_stream(seed, 9, m, 5, a, b) for m in range(4000)])
        assert scipy.stats.kstest(xs, scipy.stats.gamma(a, scale=b).cdf).pvalue > 1e-3
    xs = np.array([O.lib().or_beta_stream(seed, 0, 1, m, 7, 2.5, 7.0) for m in range(4000)])
    assert scipy.stats.kstest(xs, scipy.stats.beta(2.5, 7.0).cdf).pvalue > 1e-3


def test_uniform_slots():
    us = np.array([O.lib().or_uniform(5, c0, 3, 4, 5, lane) for c0 in range(2000) for lane in range(4)])
    assert us.min() > 0 and us.max() < 1
    assert scipy.stats.kstest(us, "uniform").pvalue > 1e-3


@pytest.mark.parametrize("case", [0, 1])
def test_z_update_matches_exact_enumeration(case):
    """Update_Z_subtree's forward-backward draw (bpp_c.cpp:1182-1293) against the
    exact distribution over every Z configuration of a 4-tip tree."""
    C = O.C
    nodes = np.ascontiguousarray(G["zc_nodes"], np.int32)
    parent = np.array(G["zc_parent"], np.int32)
    N = len(parent)
    parent[N - 1] = N
    fixZ = np.ascontiguousarray(G[f"zc{case}_fixZ"], np.int32)
    E = np.ascontiguousarray(G["zc_E"], np.float64)
    E[N - 1] = 0.0        # the root carries no branch emission
    confs, w = G[f"zc{case}_confs"], G[f"zc{case}_w"]
    index = {tuple(cf): i for i, cf in enumerate(confs)}
    counts = np.zeros(len(confs))
    Zout = np.zeros(N, np.int32)
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
    n = 20000
    for m in range(n):
        O.lib().or_test_update_Z(N, P(nodes, C.c_int), len(nodes), P(parent, C.c_int), P(fixZ, C.c_int),
                                 P(E, C.c_double), C.c_double(0.4), C.c_double(0.2), C.c_double(0.1),
                                 C.c_uint64(99), m, 1, P(Zout, C.c_int))
        counts[index[tuple(Zout[nodes])]] += 1
    exp = w * n
    keep = exp > 5
    chi2 = np.sum((counts[keep] - exp[keep]) ** 2 / exp[keep])
    dof = keep.sum() - 1
    assert scipy.stats.chi2.sf(chi2, dof) > 1e-4, (chi2, dof)
    assert counts[~keep].sum() <= max(20, 3 * exp[~keep].sum())
