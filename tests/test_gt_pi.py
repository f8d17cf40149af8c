"""PhyloAcc-GT's per-element base frequencies: the strand-symmetric model
pi = (a, 1/2 - a, 1/2 - a, a) (BPP_C::getQmat, src/PhyloAcc-GT/bpp_c.cpp:160-185,
exchangeabilities as in src/PhyloAcc-GT/bpp.hpp:346-353) and the sample_pi
Metropolis step (bpp_c.cpp:1508-1597), batched over gene-tree elements.

The oracle is pinned by what the mathematics fixes:
* a column with one observed leaf has likelihood pi of its base, whatever the
  tree (the rate matrix must have pi as its stationary distribution);
* for a model whose exchangeabilities already have the reference's tied form
  and whose pi is strand-symmetric, rebuilding the model at the same a
  changes nothing (a mistaken pair assignment breaks it);
* brute force over every ancestral assignment of tiny gene trees;
* with no data the step's stationary law is its prior: the final a of many
  independent chains follows Beta(prior)/2 (a dropped prior term gives a
  uniform law instead; the prior here keeps the chain away from the clamps
  at 0.05 and 0.45, whose boundary terms are the reference's own); the
  resampling step draws from Beta(prior + base counts)/2.
The device path is then held to the oracle: likelihoods within 1e-9 (FP64)
and 1e-6 (FP32), and identical step trajectories.
"""
import numpy as np
import pytest
import scipy.linalg
import scipy.stats

from oracle import oracle_py as O
from phyloacc_amd import engine as E
from phyloacc_amd import synth
from test_gene_trees import LEAF, _tiny

SEED = 7


def _strand_model(a, r=(1.3, 4.1, 0.7)):
    """A model already in the tied form: exchangeabilities AC = CG = CT = r0,
    AG = GT = r1, AT = r2, and pi = (a, 1/2 - a, 1/2 - a, a)."""
    pi = np.array([a, 0.5 - a, 0.5 - a, a])
    R = np.zeros((4, 4))
    for (i, j), v in {(0, 1): r[0], (0, 2): r[1], (0, 3): r[2], (1, 2): r[0], (1, 3): r[0], (2, 3): r[1]}.items():
        R[i, j] = R[j, i] = v
    Q = R * pi[None, :]
    np.fill_diagonal(Q, -Q.sum(1))
    return pi, Q


def _gt_call(fn, m, tree, gt, Z, align, *extra):
    return fn(m, *extra, tree.S, gt.children, gt.seg_gene, gt.seg_species, gt.seg_len, Z, 0.3, 2.0, align, 0,
              align.shape[1])


def test_single_observed_leaf_gives_pi():
    rng = np.random.default_rng(1)
    tree, h, gt, pi, Q, _ = _tiny(1, S=5)
    m, _ = O.parse_mod(synth.mod_text(pi, Q, tree))
    Z = np.ones(tree.N, np.int32)
    for a in (0.12, 0.3, 0.41):
        for leaf in range(tree.S):
            align = np.full((tree.S, 4), ord("-"), np.uint8)
            align[leaf] = np.frombuffer(b"acgt", np.uint8)
            got = _gt_call(O.gt_loglik_pi, m, tree, gt, Z, align, a)
            want = np.log(a) + np.log(0.5 - a) + np.log(0.5 - a) + np.log(a)
            assert abs(got - want) < 1e-12, (a, leaf, got, want)


def test_tied_model_rebuilds_to_itself():
    tree, h, gt, _, _, align = _tiny(2, S=5, GG=20)
    Z = np.array([1, 2, 1, 1, 2, 1, 1, 2, 1], np.int32)
    for a in (0.18, 0.27):
        pi, Q = _strand_model(a)
        m, _ = O.parse_mod(synth.mod_text(pi, Q, tree))
        base = _gt_call(O.gt_loglik, m, tree, gt, Z, align)
        assert abs(_gt_call(O.gt_loglik_pi, m, tree, gt, Z, align, a) - base) <= 1e-12 * abs(base)
        assert abs(_gt_call(O.gt_loglik_pi, m, tree, gt, Z, align, a + 0.05) - base) > 1e-6


@pytest.mark.parametrize("seed", [3, 4])
def test_oracle_pi_brute_force(seed):
    """The rebuilt model (getQmat at pi(a) with the tied exchangeabilities
    taken from the model) against a brute-force sum with scipy's expm."""
    tree, h, gt, pi0, Q0, align = _tiny(seed, S=4, GG=10)
    m, _ = O.parse_mod(synth.mod_text(pi0, Q0, tree))
    Qm, pm = np.array(m.Q[:16]).reshape(4, 4), np.array(m.pi[:4])
    inst = [Qm[0, 1] / pm[1], Qm[0, 2] / pm[2], Qm[0, 3] / pm[3]]
    inst = [inst[0], inst[1], inst[2], inst[0], Qm[1, 2] / pm[2], inst[1]]
    a = 0.21
    pi = np.array([a, 0.5 - a, 0.5 - a, a])
    R = np.zeros((4, 4))
    for k, (i, j) in enumerate([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]):
        R[i, j] = R[j, i] = inst[k]
    Q = R * pi[None, :]
    np.fill_diagonal(Q, -Q.sum(1))
    Z = np.ones(tree.N, np.int32)
    want = _brute_pi(gt, pi, Q, align, Z, 0.3)
    got = _gt_call(O.gt_loglik_pi, m, tree, gt, Z, align, a)
    assert abs(got - want) <= 1e-10 * abs(want), (got, want)


def _brute_pi(gt, pi, Q, align, Z, r1):
    import itertools
    S = align.shape[0]
    N = 2 * S - 1
    P = [np.eye(4) for _ in range(N)]
    for g, s, L in zip(gt.seg_gene, gt.seg_species, gt.seg_len):
        if L < 1e-8:
            continue
        P[g] = P[g] @ scipy.linalg.expm(Q * L * (r1 if Z[s] == 1 else (2.0 if Z[s] == 2 else 1.0)))
    total = 0.0
    for col in range(align.shape[1]):
        leaf = [np.array(LEAF[int(align[s, col])], float) for s in range(S)]
        acc = 0.0
        for states in itertools.product(range(4), repeat=N - S):
            st = dict(zip(range(S, N), states))
            p = pi[st[N - 1]]
            for v in range(N - 1):
                par = st[gt.parent[v]]
                p *= (P[v][par] @ leaf[v]) if v < S else P[v][par, st[v]]
            acc += p
        total += np.log(acc)
    return total


def test_sample_pi_prior_recovery_and_resampling():
    tree, h, gt, pi, Q, _ = _tiny(5, S=4)
    m, _ = O.parse_mod(synth.mod_text(pi, Q, tree))
    Z = np.ones(tree.N, np.int32)
    gaps = np.full((tree.S, 3), ord("-"), np.uint8)           # no data: log-likelihood 0
    prior = (20.0, 20.0)
    finals = []
    for chain in range(300):
        a, ll = 0.25, 0.0
        for sweep in range(60):
            a, ll, _ = _gt_call(O.gt_sample_pi, m, tree, gt, Z, gaps, 1000 + chain, sweep, 0.05, prior, 0, a, ll)
        finals.append(a)
    ks = scipy.stats.kstest(2 * np.array(finals), scipy.stats.beta(*prior).cdf)
    assert ks.pvalue > 1e-3, ks
    # resampling: a = Beta(prior + #A + #T, prior + #C + #G) / 2 over the element's leaf bases
    data = np.frombuffer(b"aacgttgcaa", np.uint8)
    align = np.tile(data, (tree.S, 1))
    nat = tree.S * 6
    ncg = tree.S * 4
    draws = [_gt_call(O.gt_sample_pi, m, tree, gt, Z, align, SEED, sweep, 0.05, (2.0, 3.0), 1, 0.25, 0.0)[0]
             for sweep in range(400)]
    ks = scipy.stats.kstest(2 * np.array(draws), scipy.stats.beta(2.0 + nat, 3.0 + ncg).cdf)
    assert ks.pvalue > 1e-3, ks


def _gt_engine(seed=11, n=12, S=30):
    w = synth.make_gt_workload(seed=seed, n_elements=n, S=S)
    opts = E.default_options(min_len=1, seed=SEED)
    eng = E.Engine.from_workload(w, opts)
    ids, st = eng.add_elements_with_trees(w.align, w.starts, w.sizes, w.children())
    assert (st == E.PA_ELEM_OK).all()
    return w, eng, ids, O.parse_mod(w.mod_text)[0], opts


def _oracle_args(w, c):
    g = w.genes[c]
    return (w.tree.S, g.children, g.seg_gene, g.seg_species, g.seg_len * w.scale, w.Z[c], w.r1[c], w.r2[c],
            w.align, w.starts[c], w.sizes[c])


@pytest.mark.gpu
def test_loglik_with_element_pi_vs_oracle(require_gpu):
    w, eng, ids, m, opts = _gt_engine()
    rng = np.random.default_rng(2)
    a = rng.uniform(0.08, 0.42, len(ids))
    L = w.lengths()
    ref = np.array([O.gt_loglik_pi(m, a[c], *_oracle_args(w, c), gap_prop=opts.gap_prop) for c in range(len(ids))])
    got64 = eng.loglik_lengths_pi(ids, L, a, E.PA_FP64)
    got32 = eng.loglik_lengths_pi(ids, L, a, E.PA_FP32)
    np.testing.assert_allclose(got64, ref, rtol=1e-9)
    np.testing.assert_allclose(got32, ref, rtol=1e-6)
    with pytest.raises(E.PaError):
        eng.loglik_lengths_pi(ids, L, np.full(len(ids), 0.5))


@pytest.mark.gpu
def test_sample_pi_steps_match_oracle(require_gpu):
    """Batched steps against the oracle's, element by element: the same a
    trajectory and decisions, log-likelihoods within 1e-9 (FP64); one
    resampling step in the middle, as the reference's Gibbs takes one when the
    acceptance rate is low (bpp_c.cpp:543-555)."""
    w, eng, ids, m, opts = _gt_engine(seed=12, n=10)
    L = w.lengths()
    a = np.full(len(ids), 0.25)
    ll = eng.loglik_lengths_pi(ids, L, a, E.PA_FP64)
    oa, oll = a.copy(), np.array([O.gt_loglik_pi(m, 0.25, *_oracle_args(w, c), gap_prop=opts.gap_prop)
                                  for c in range(len(ids))])
    np.testing.assert_allclose(ll, oll, rtol=1e-9)
    prior = (2.0, 2.0)
    n_acc = 0
    for sweep in range(12):
        resample = sweep == 6
        a, ll, acc = eng.sample_pi(ids, L, sweep, a, ll, pi_delta=0.05, prior=prior, resample=resample)
        for c in range(len(ids)):
            oa[c], oll[c], oacc = O.gt_sample_pi(m, SEED, sweep, 0.05, prior, resample, oa[c], oll[c],
                                                 *_oracle_args(w, c), gap_prop=opts.gap_prop)
            assert acc[c] == oacc, (sweep, c)
        np.testing.assert_array_equal(a, oa)
        np.testing.assert_allclose(ll, oll, rtol=1e-9)
        n_acc += int(acc.sum())
    assert 0 < n_acc < 12 * len(ids)              # both outcomes occur
