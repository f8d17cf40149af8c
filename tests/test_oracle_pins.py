"""Pins the oracle's samplers and Eval2 against what the mathematics fixes.

These are the oracle functions the brute-force likelihood checks leave open:

* the Metropolis updates of r2 and r1 (prior terms, Gamma-proposal Hastings
  terms, the multiplicative proposal's Jacobian, the ROPT truncations) and the
  Beta updates of g, l, l2 (src/PhyloAcc-ST/bpp_c.cpp:522-613, 930-1176): with
  no data the chain's stationary distribution is the prior, so the final states
  of many independent chains must follow the Gamma and Beta priors (KS tests);
* the ancestral draw (bpp_c.cpp:672-733): the joint draw of the internal bases
  of one site must follow the posterior given the leaves, enumerated by brute
  force with scipy's expm (chi-square test);
* Eval2 (src/PhyloAcc-ST/bpp_c2.cpp:14-353): on traces whose configurations
  all carry weight, it is the variational bound
      sum_i w_i (E_i[log p(X, r)] + log P(Z_i) + 1/2 log det(2 pi e Sigma_i)) + H(w),
  and when no configuration does, log sum_i P(X | Z_i, r*) P(Z_i)
  + log p(r*) + 1/2 log det(2 pi Sigma); here log P(Z) is integrated
  numerically over the Beta hyperpriors and P(X | Z, r*) comes from the
  brute-force-pinned likelihood.

A plausible mistake -- a dropped Hastings term, a swapped transition count, a
draw from the wrong conditional, a missing entropy term -- moves one of these
away from its target.
"""
import itertools
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import scipy.integrate
import scipy.linalg
import scipy.special
import scipy.stats

from oracle import oracle_py as O
from phyloacc_amd import engine as E
from phyloacc_amd import synth

NT = max(1, min(8, os.cpu_count() or 1))


# ------------------------------------------------------------ prior recovery

# No data: 100 columns that are gaps in every species (loglik 0 at any Z and
# rates; 100 columns make the r2 proposal scale 10 / GG usable).  A
# HYPER_LRATE prior of Beta(50, 1) keeps accelerated branches present in
# almost every sweep, so r2 moves by its Metropolis update rather than by a
# draw from its prior; adaptation is off and r1 draws a fresh uniform (the
# reference's shared uniform, r1_shared_uniform = 1, does not leave the prior
# invariant).
NO_DATA = dict(min_len=1, gap_prop=1.0, conserve_prop=1.0, r1_shared_uniform=0, adapt_freq=10 ** 6,
               hyper_lrate_a=50.0, hyper_lrate_b=1.0)


def _no_data_chains(ropt, n_chains, T, length=100):
    """Final states of n_chains independent three-model chains on a no-data element."""
    w = synth.small_workload(S=5, C=1, length=length, seed=12)
    w.align[:] = ord("-")
    m = O.model_from_workload(w)
    kw = dict(burnin=T // 2, mcmc=T - T // 2, ropt=ropt, **NO_DATA)

    def one(seed):
        oo = O.default_options(seed=seed, **kw)
        e = O.Element(m, oo, w.align, 0, length)
        assert not e.filtered
        r = e.run(oo)
        return [(r[ms]["nrate"][-1], r[ms]["crate"][-1], r[ms]["grate"][-1], r[ms]["lrate"][-1], r[ms]["l2rate"][-1])
                for ms in range(3)]

    with ThreadPoolExecutor(NT) as ex:
        out = list(ex.map(one, range(1, n_chains + 1)))
    return np.array(out)      # [chain, model, (r2, r1, g, l, l2)]


def _truncated(dist, lo=None, hi=None):
    flo = dist.cdf(lo) if lo is not None else 0.0
    fhi = dist.cdf(hi) if hi is not None else 1.0
    return lambda x: np.clip((dist.cdf(x) - flo) / (fhi - flo), 0, 1)


@pytest.mark.parametrize("ropt", [0, 1])
def test_no_data_chain_recovers_the_priors(ropt):
    o = O.default_options(**NO_DATA)
    x = _no_data_chains(ropt, n_chains=600, T=400)
    r2 = x[:, 2, 0]     # M2 (full model): r2 moves by Metropolis or by a prior draw
    r1 = x[:, 2, 1]
    g, l, l2 = x[:, 2, 2], x[:, 2, 3], x[:, 2, 4]
    prior_r2 = scipy.stats.gamma(o.nprior_a, scale=o.nprior_b)
    prior_r1 = scipy.stats.gamma(o.cprior_a, scale=o.cprior_b)
    if ropt == 1:
        cdf_r2 = _truncated(prior_r2, lo=o.nlb)     # r2 >= NLB
        cdf_r1 = _truncated(prior_r1, hi=o.cub)     # r1 <= CUB
        assert r2.min() >= o.nlb and r1.max() <= o.cub
    else:
        cdf_r2, cdf_r1 = prior_r2.cdf, prior_r1.cdf
    tests = {
        "r2": scipy.stats.kstest(r2, cdf_r2).pvalue,
        "r1": scipy.stats.kstest(r1, cdf_r1).pvalue,
        "g": scipy.stats.kstest(g, scipy.stats.beta(o.hyper_grate_a, o.hyper_grate_b).cdf).pvalue,
        "l": scipy.stats.kstest(l, scipy.stats.beta(o.hyper_lrate_a, o.hyper_lrate_b).cdf).pvalue,
        "l2": scipy.stats.kstest(l2, scipy.stats.beta(o.hyper_lrate2_a, o.hyper_lrate2_b).cdf).pvalue,
    }
    print("prior recovery KS p-values:", tests)
    assert min(tests.values()) > 1e-4, tests
    # the null model never samples r2 (resZ 0 keeps it at INIT_ACCE_RATE)
    assert np.all(x[:, 0, 0] == o.init_nrate)


@pytest.mark.parametrize("mutant,param", [("OR_MUTANT_NO_HASTINGS", "r2"), ("OR_MUTANT_SWAP_G_COUNTS", "g")])
def test_prior_recovery_catches_planted_mistakes(tmp_path, mutant, param):
    """The prior-recovery check on copies of the oracle with a planted
    mistake (compiled here, on the CPU): the Gamma proposal's Hastings terms
    dropped, or g's Beta counts swapped.  The check must reject both."""
    src = os.path.join(os.path.dirname(O.__file__), "pa_oracle.c")
    so = str(tmp_path / f"mutant_{mutant}.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-std=c11", "-D_GNU_SOURCE", f"-D{mutant}", "-shared", "-o", so,
                    src, "-lm"], check=True)
    saved = O._lib
    O._lib = O.load(so)
    try:
        x = _no_data_chains(0, n_chains=600, T=400)
    finally:
        O._lib = saved
    o = O.default_options(**NO_DATA)
    if param == "r2":
        p = scipy.stats.kstest(x[:, 2, 0], scipy.stats.gamma(o.nprior_a, scale=o.nprior_b).cdf).pvalue
    else:
        p = scipy.stats.kstest(x[:, 2, 2], scipy.stats.beta(o.hyper_grate_a, o.hyper_grate_b).cdf).pvalue
    print(f"{mutant}: KS p-value of {param} against its prior {p:.3g}")
    assert p < 1e-4


# --------------------------------------------------------- ancestral draws

def test_ancestral_draw_matches_the_brute_force_posterior():
    w = synth.small_workload(S=4, C=1, length=1, seed=21, mean_bl=0.6)
    w.align[:, 0] = np.frombuffer(b"acca", np.uint8)
    m = O.model_from_workload(w)
    oo = O.default_options(min_len=1)
    e = O.Element(m, oo, w.align, 0, 1)
    a = e.arrays()
    S, N = a["S"], a["N"]
    assert a["n_internal"] == 3
    Z = np.ones(N, np.int32)
    Z[N - 1] = 0
    Z[a["nodes"][0]] = 2
    r1, r2 = 0.7, 2.5
    rate = {0: 1.0, 1: r1, 2: r2}
    pi = np.array(m.pi[:4])
    Q = np.array(m.Q[:16]).reshape(4, 4)
    P = {v: scipy.linalg.expm(Q * a["dist"][v] * rate[int(Z[v])]) for v in range(N - 1)}
    leaf = {s: "acgt".index(chr(w.align[s, 0])) for s in range(S)}
    internal = list(range(S, N))
    post = {}
    for states in itertools.product(range(4), repeat=len(internal)):
        x = dict(zip(internal, states))
        x.update(leaf)
        p = pi[x[N - 1]]
        for v in range(N - 1):
            p *= P[v][x[a["parent"][v]], x[v]]
        post[states] = p
    tot = sum(post.values())
    n = 30000
    counts = dict.fromkeys(post, 0)
    for mm in range(n):
        st = e.sample_site(Z, r1, r2, seed=17, m_sweep=mm)
        counts[tuple(int(st[v]) for v in internal)] += 1
    exp = np.array([post[k] / tot * n for k in post])
    obs = np.array([counts[k] for k in post])
    keep = exp >= 5
    chi2 = np.sum((obs[keep] - exp[keep]) ** 2 / exp[keep])
    pval = scipy.stats.chi2.sf(chi2, keep.sum() - 1)
    print(f"ancestral draw chi2 {chi2:.1f} on {keep.sum() - 1} dof, p = {pval:.3g}")
    assert pval > 1e-4
    assert obs[~keep].sum() <= max(20, 3 * exp[~keep].sum())


# ------------------------------------------------------------------- Eval2

def _element(seed=9, S=6, length=40):
    w = synth.small_workload(S=S, C=1, length=length, seed=seed, gap=0.0, mean_bl=0.3)
    m = O.model_from_workload(w)
    return w, m


def _fixZ(m, a, ms):
    """fixZ of model slot ms (BPP_C::initMCMC, bpp_c.cpp:240-289): the
    outgroup lineage and the root; M1 adds the CONSERVE ancestors; M0 every
    branch."""
    N, root = a["N"], a["nodes"][-1]
    fz = np.zeros(N, bool)
    for s in a["nodes"]:
        fz[s] = bool(m.upper[s]) or s == root or ms == 0 or (ms == 1 and bool(m.upper_conserve[s]))
    return fz


def _log_prior_Z(Zc, a, fz, o):
    """log P(Z) under the transition model with Beta hyperpriors, integrated
    numerically over g, l and l2 (root state prior 1/2, 1/2, 0)."""
    nodes, par = a["nodes"], a["parent"]
    n_all = np.zeros((3, 3))
    n_free = np.zeros((3, 3))
    for s in nodes[:-1]:
        n_all[Zc[s], Zc[par[s]]] += 1
        if not fz[s]:
            n_free[Zc[s], Zc[par[s]]] += 1

    def integral(a_, b_, k_succ, k_fail):
        if a_ == 0:          # Dollo: l2 = 0, the factor is 1 when no branch uses it
            return 0.0 if k_succ + k_fail == 0 else -np.inf
        f = lambda x: scipy.stats.beta.pdf(x, a_, b_) * x ** k_succ * (1 - x) ** k_fail
        val, _ = scipy.integrate.quad(f, 0, 1, epsabs=0, epsrel=1e-12, limit=200)
        return np.log(val)

    lp = np.log(0.5)
    lp += integral(o.hyper_grate_a, o.hyper_grate_b, n_all[1, 0] + n_all[2, 0], n_all[0, 0])
    lp += integral(o.hyper_lrate_a, o.hyper_lrate_b, n_free[2, 1], n_free[1, 1])
    lp += integral(o.hyper_lrate2_a, o.hyper_lrate2_b, n_free[1, 2], n_free[2, 2])
    return lp


def _configs(a, fz, k, rng):
    """k distinct valid Z configurations: root 1, fixZ branches 1, free
    branches 1 or 2 (every transition has positive probability)."""
    N, nodes = a["N"], a["nodes"]
    free = [s for s in nodes[:-1] if not fz[s]]
    assert 2 ** len(free) >= k
    picks = rng.choice(2 ** len(free), size=k, replace=False)
    out = []
    for p in picks:
        Zc = np.full(N, -1, np.int8)
        for s in nodes:
            Zc[s] = 1
        for j, s in enumerate(free):
            Zc[s] = 2 if (p >> j) & 1 else 1
        out.append(Zc)
    return out


def _traces(configs, weights, nb, nm, rng):
    """Synthetic traces: burn-in rows at the first configuration, then nm rows
    holding configuration i for round(weights[i] * nm) rows; rates and full
    log-likelihoods drawn per configuration."""
    reps = np.round(np.asarray(weights) * nm).astype(int)
    reps[-1] = nm - reps[:-1].sum()
    rows = [0] * nb + [i for i, r in enumerate(reps) for _ in range(r)]
    T = nb + nm
    Z = np.stack([configs[i] for i in rows])
    crate = np.array([rng.uniform(0.1, 0.5) + 0.05 * rows[t] for t in range(T)])
    nrate = np.array([rng.uniform(1.0, 3.0) + 0.3 * crate[t] for t in range(T)])
    full = np.array([rng.normal(-200.0 - 3 * rows[t], 2.0) for t in range(T)])
    return Z, crate, nrate, full, np.array(rows)


def _elbo(configs, rows, Z, crate, nrate, full, nb, nm, a, fz, o, resZ):
    """sum_i w_i (E_i[full] + log P(Z_i) + 1/2 log det(2 pi e Sigma_i)) + H(w)."""
    post = rows[nb:]
    tot = 0.0
    for i in np.unique(post):
        sel = np.flatnonzero(post == i) + nb
        wi = len(sel) / nm
        X = np.stack([crate[sel], nrate[sel]]) if resZ != 0 else crate[sel][None, :]
        Sig = np.atleast_2d(np.cov(X, bias=True))
        d = Sig.shape[0]
        ent = 0.5 * np.linalg.slogdet(Sig)[1] + 0.5 * d * (np.log(2 * np.pi) + 1)
        tot += wi * (full[sel].mean() + _log_prior_Z(configs[i], a, fz, o) + ent) - wi * np.log(wi)
    return tot


@pytest.mark.parametrize("ms,weights", [(2, [1.0]), (2, [0.6, 0.3, 0.1]), (0, [1.0]), (1, [0.7, 0.3])])
def test_eval2_is_the_variational_bound(ms, weights):
    w, m = _element()
    nb, nm = 20, 100
    o = O.default_options(burnin=nb, mcmc=nm, min_len=1)
    e = O.Element(m, o, w.align, 0, w.sizes[0])
    a = e.arrays()
    fz = _fixZ(m, a, ms)
    rng = np.random.default_rng(ms * 10 + len(weights))
    resZ = (0, 2, 1)[ms]
    if ms == 0:
        configs = [np.where(np.arange(a["N"]) >= 0, 1, 1).astype(np.int8)]
        configs[0][[s for s in range(a["N"]) if s not in set(a["nodes"])]] = -1
    else:
        configs = _configs(a, fz, len(weights), rng)
    Z, crate, nrate, full, rows = _traces(configs, weights, nb, nm, rng)
    ref = _elbo(configs, rows, Z, crate, nrate, full, nb, nm, a, fz, o, resZ)
    maxZ = configs[0].astype(np.int32)
    vb_oracle = e.eval2(o, ms, full, nrate, crate, Z, maxZ)
    print(f"model {ms} weights {weights}: oracle {vb_oracle:.10f} bound {ref:.10f}")
    assert abs(vb_oracle - ref) <= 1e-9 * abs(ref)
    # the engine's host summary (pa_chain_summary) on the same traces
    eng = E.Engine.from_workload(w, E.default_options(burnin=nb, mcmc=nm, min_len=1))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    tr = dict(full=full, nrate=nrate, crate=crate, grate=np.zeros(nb + nm), lrate=np.zeros(nb + nm),
              l2rate=np.zeros(nb + nm), Z=Z, max_Z=configs[0])
    got = eng.chain_summary(int(ids[0]), ms, tr)["vb"]
    assert abs(got - ref) <= 1e-9 * abs(ref), (got, ref)


def test_eval2_fallback_is_the_marginal_over_configurations():
    """No configuration reaches 2% of the states: Eval2 falls back to
    log sum_i P(X | Z_i, r*) P(Z_i) + log p(r*) + 1/2 log det(2 pi Sigma),
    r* the posterior mean rates (bpp_c2.cpp:198-297)."""
    w, m = _element(seed=5, S=8)
    nb, nm = 10, 60
    o = O.default_options(burnin=nb, mcmc=nm, min_len=1)
    e = O.Element(m, o, w.align, 0, w.sizes[0])
    a = e.arrays()
    ms, resZ = 2, 1
    fz = _fixZ(m, a, ms)
    rng = np.random.default_rng(8)
    configs = _configs(a, fz, nm, rng)
    Z, crate, nrate, full, rows = _traces(configs, [1.0 / nm] * nm, nb, nm, rng)
    vb = e.eval2(o, ms, full, nrate, crate, Z, configs[0].astype(np.int32))
    rc, rn = crate[nb:].mean(), nrate[nb:].mean()
    terms = []
    for i in np.unique(rows[nb:]):
        ll = O.Element(m, o, w.align, 0, w.sizes[0]).loglik(configs[i].astype(np.int32), rc, rn)
        terms.append(ll + _log_prior_Z(configs[i], a, fz, o))
    Sig = np.cov(np.stack([crate[nb:], nrate[nb:]]), bias=True)
    ref = (scipy.special.logsumexp(terms) + scipy.stats.gamma.logpdf(rc, o.cprior_a, scale=o.cprior_b)
           + scipy.stats.gamma.logpdf(rn, o.nprior_a, scale=o.nprior_b)
           + 0.5 * np.linalg.slogdet(Sig)[1] + np.log(2 * np.pi))
    print(f"fallback: oracle {vb:.10f} marginal {ref:.10f}")
    assert abs(vb - ref) <= 1e-9 * abs(ref)


@pytest.mark.gpu
def test_engine_eval2_fallback_matches_the_marginal(require_gpu):
    """The engine's Eval2 fallback re-evaluates P(X | Z_i, r*) with pa_loglik
    on the device; it gives the same marginal as the oracle and the bound."""
    w, m = _element(seed=5, S=8)
    nb, nm = 10, 60
    o = O.default_options(burnin=nb, mcmc=nm, min_len=1)
    e = O.Element(m, o, w.align, 0, w.sizes[0])
    a = e.arrays()
    fz = _fixZ(m, a, 2)
    rng = np.random.default_rng(8)
    configs = _configs(a, fz, nm, rng)
    Z, crate, nrate, full, rows = _traces(configs, [1.0 / nm] * nm, nb, nm, rng)
    ref = e.eval2(o, 2, full, nrate, crate, Z, configs[0].astype(np.int32))
    eng = E.Engine.from_workload(w, E.default_options(burnin=nb, mcmc=nm, min_len=1))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    tr = dict(full=full, nrate=nrate, crate=crate, grate=np.zeros(nb + nm), lrate=np.zeros(nb + nm),
              l2rate=np.zeros(nb + nm), Z=Z, max_Z=configs[0])
    got = eng.chain_summary(int(ids[0]), 2, tr, precision=E.PA_FP64)["vb"]
    assert abs(got - ref) <= 1e-9 * abs(ref), (got, ref)
