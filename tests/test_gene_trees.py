"""Gene-tree pruning (PhyloAcc-GT style): elements that carry their own tree.

The oracle's gene-tree likelihood (one transition matrix per piece of a gene
branch inside a species branch, composed in log space) is pinned against a
brute-force sum over every ancestral assignment with scipy's expm, and
against the species-tree likelihood when the gene tree equals the species
tree.  The engine's gene lengths, encoding and (on the GPU) batched pruning
are then checked against the oracle.
"""
import itertools

import numpy as np
import pytest
import scipy.linalg

from oracle import oracle_py as O
from phyloacc_amd import engine as E
from phyloacc_amd import synth


def _tiny(seed, S=4, GG=12, gap=0.1):
    rng = np.random.default_rng(seed)
    tree, h = synth.ultrametric_tree(S, 0.4, rng)
    gt = synth.msc_gene_tree(tree, h, rng)
    pi, Q = synth.gtr(rng)
    align = rng.choice(np.frombuffer(b"acgt-nr", np.uint8), size=(S, GG),
                       p=[(1 - gap) / 4] * 4 + [gap / 3] * 3)
    return tree, h, gt, pi, Q, np.ascontiguousarray(align)


LEAF = {ord("a"): [1, 0, 0, 0], ord("c"): [0, 1, 0, 0], ord("g"): [0, 0, 1, 0], ord("t"): [0, 0, 0, 1],
        ord("r"): [1, 0, 1, 0], ord("-"): [1, 1, 1, 1], ord("n"): [1, 1, 1, 1]}


def _brute(gt, pi, Q, align, Z, r1, r2, scale=1.0):
    """Sum over all 4^I internal assignments; each gene branch's matrix is the
    product of expm over its pieces (rate by the piece's species-branch Z)."""
    S = align.shape[0]
    N = 2 * S - 1
    rate = {0: 1.0, 1: r1, 2: r2}
    P = [np.eye(4) for _ in range(N)]
    for g, s, L in zip(gt.seg_gene, gt.seg_species, gt.seg_len):
        if L * scale < 1e-8:
            continue
        P[g] = P[g] @ scipy.linalg.expm(Q * L * scale * rate[int(Z[s])])
    total = 0.0
    for col in range(align.shape[1]):
        leaf = [np.array(LEAF[int(align[s, col])], float) for s in range(S)]
        acc = 0.0
        for states in itertools.product(range(4), repeat=N - S):
            st = dict(zip(range(S, N), states))
            p = pi[st[N - 1]]
            for v in range(N - 1):
                a = st[gt.parent[v]]
                if v < S:
                    p *= P[v][a] @ leaf[v]
                else:
                    p *= P[v][a, st[v]]
            acc += p
        total += np.log(acc)
    return total


def _model(tree, pi, Q):
    m, _ = O.parse_mod(synth.mod_text(pi, Q, tree))
    return m


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_gene_tree_brute_force(seed):
    tree, h, gt, pi, Q, align = _tiny(seed)
    m = _model(tree, pi, Q)
    rng = np.random.default_rng(seed + 10)
    Z = rng.integers(0, 3, size=tree.N).astype(np.int32)
    ref = _brute(gt, np.array(m.pi[:4]), np.array(m.Q[:16]).reshape(4, 4), align, Z, 0.3, 2.2)
    ll = O.gt_loglik(m, tree.S, gt.children, gt.seg_gene, gt.seg_species, gt.seg_len, Z, 0.3, 2.2,
                     align, 0, align.shape[1], gap_prop=1.0)
    assert abs(ll - ref) <= 1e-11 * abs(ref), (ll, ref)


def test_oracle_gene_tree_equal_to_species_tree():
    # a gene tree identical to the species tree, one piece per branch, gives the
    # species-tree likelihood (bpp_c2.cpp:392-425 form) at the same Z and rates
    w = synth.small_workload(S=6, C=1, length=50, seed=9, gap=0.03)
    m = O.model_from_workload(w)
    t = w.tree
    seg_gene = np.arange(t.N - 1, dtype=np.int32)
    Z = np.random.default_rng(2).integers(0, 3, size=t.N).astype(np.int32)
    Z[t.N - 1] = 1
    gl = O.gt_loglik(m, t.S, t.children, seg_gene, seg_gene, t.dist[: t.N - 1], Z, 0.35, 1.9,
                     w.align, 0, 50, gap_prop=1.0)
    opts = O.default_options(min_len=1, gap_prop=1.0)
    st = O.Element(m, opts, w.align, 0, 50).loglik(Z, 0.35, 1.9)
    assert abs(gl - st) <= 1e-11 * abs(st), (gl, st)


def test_gene_lengths_match_pieces():
    w = synth.make_gt_workload(seed=3, n_elements=12, S=20)
    for c in range(w.C):
        g = w.genes[c]
        out = E.gene_lengths(w.tree.N, g.seg_gene, g.seg_species, g.seg_len * w.scale, w.Z[c], w.r1[c], w.r2[c])
        np.testing.assert_allclose(out, w.lengths()[c], rtol=1e-14, atol=1e-300)
        # at rate 1 everywhere, each gene branch spans its two heights
        ones = E.gene_lengths(w.tree.N, g.seg_gene, g.seg_species, g.seg_len, np.zeros(w.tree.N, np.int32), 1, 1)
        span = np.array([g.height[g.parent[v]] - g.height[v] for v in range(w.tree.N - 1)])
        leaf_span = np.array([g.height[g.parent[v]] - w.sp_height[v] for v in range(w.tree.S)])
        span[: w.tree.S] = leaf_span
        np.testing.assert_allclose(ones[:-1], span, rtol=1e-9, atol=1e-12)


def test_add_trees_encoding():
    w = synth.make_gt_workload(seed=4, n_elements=6, S=12)
    eng = E.Engine.from_workload(w, E.default_options(min_len=1))
    ids, st = eng.add_elements_with_trees(w.align, w.starts, w.sizes, w.children())
    assert (st == E.PA_ELEM_OK).all()
    for c, i in enumerate(ids):
        el = eng.element(int(i))
        np.testing.assert_array_equal(el["children"], w.genes[c].children)
        assert el["root"] == w.tree.N - 1
    bad = w.children().copy()
    bad[0, w.tree.N - 1] = [w.tree.N - 1, 0]
    with pytest.raises(E.PaError):
        eng.add_elements_with_trees(w.align, w.starts, w.sizes, bad)


@pytest.mark.gpu
def test_gene_tree_loglik_gpu_vs_oracle(require_gpu):
    w = synth.make_gt_workload(seed=5, n_elements=48, S=40)
    opts = E.default_options(min_len=1)
    eng = E.Engine.from_workload(w, opts)
    ids, st = eng.add_elements_with_trees(w.align, w.starts, w.sizes, w.children())
    assert (st == E.PA_ELEM_OK).all()
    m = O.parse_mod(w.mod_text)[0]
    L = w.lengths()
    ref = np.array([O.gt_loglik(m, w.tree.S, g.children, g.seg_gene, g.seg_species, g.seg_len * w.scale, w.Z[c],
                                w.r1[c], w.r2[c], w.align, w.starts[c], w.sizes[c], gap_prop=opts.gap_prop)
                    for c, g in enumerate(w.genes)])
    got64 = eng.loglik_lengths(ids, L, E.PA_FP64)
    np.testing.assert_allclose(got64, ref, rtol=1e-9)
    got32 = eng.loglik_lengths(ids, L, E.PA_FP32)
    np.testing.assert_allclose(got32, ref, rtol=1e-6)


@pytest.mark.gpu
def test_gene_tree_loglik_gpu_vs_oracle_120_species(require_gpu):
    """The gene-tree configuration's own shape: 120-leaf gene trees (239
    nodes, 119 internal), each element with its own multispecies-coalescent
    topology."""
    w = synth.make_gt_workload(seed=8, n_elements=16, S=120)
    opts = E.default_options(min_len=1)
    eng = E.Engine.from_workload(w, opts)
    ids, st = eng.add_elements_with_trees(w.align, w.starts, w.sizes, w.children())
    assert (st == E.PA_ELEM_OK).all()
    assert len({tuple(map(tuple, g.children)) for g in w.genes}) == len(w.genes)   # shapes differ
    m = O.parse_mod(w.mod_text)[0]
    L = w.lengths()
    ref = np.array([O.gt_loglik(m, w.tree.S, g.children, g.seg_gene, g.seg_species, g.seg_len * w.scale, w.Z[c],
                                w.r1[c], w.r2[c], w.align, w.starts[c], w.sizes[c], gap_prop=opts.gap_prop)
                    for c, g in enumerate(w.genes)])
    got64 = eng.loglik_lengths(ids, L, E.PA_FP64)
    got32 = eng.loglik_lengths(ids, L, E.PA_FP32)
    print("S=120 gene trees rel err: fp64 %.2e fp32 %.2e"
          % (np.max(np.abs(got64 - ref) / np.abs(ref)), np.max(np.abs(got32 - ref) / np.abs(ref))))
    np.testing.assert_allclose(got64, ref, rtol=1e-9)
    np.testing.assert_allclose(got32, ref, rtol=1e-6)


def _counted_bytes(eng, ids, word):
    """Bytes a full pruning pass must move, from each element's tree: for
    every internal node that is not a cherry (a node whose two children are
    leaves is recomputed by its parent), one partial + exponent written and,
    per child, a partial + exponent, a leaf code or a cherry's two codes
    read; at the root, its partial + exponent and the pattern weight."""
    total = 0
    for i in ids:
        d = eng.element(int(i))
        ch, S, P, root = d["children"], eng.S, d["P"], d["root"]
        internal = [v for v in d["nodes"] if v >= S]
        cherry = {v for v in internal if ch[v][0] < S and ch[v][1] < S and v != root and len(internal) > 1}
        for v in internal:
            if v in cherry:
                continue
            per = word + 4
            for x in ch[v]:
                per += 1 if x < S else (2 if x in cherry else word + 4)
            total += P * per
        total += P * (word + 8)
    return total


@pytest.mark.gpu
@pytest.mark.parametrize("precision,word", [(E.PA_FP32, 16), (E.PA_FP64, 32)])
def test_loglik_counted_bytes(require_gpu, precision, word):
    """The likelihood kernel's on-device byte count (the numerator of its
    roofline figure) equals the count from the elements' trees."""
    w = synth.make_gt_workload(seed=10, n_elements=12, S=30)
    eng = E.Engine.from_workload(w, E.default_options(min_len=1))
    ids, st = eng.add_elements_with_trees(w.align, w.starts, w.sizes, w.children())
    eng.loglik_lengths(ids, w.lengths(), precision)
    stats = eng.loglik_stats()
    assert stats["algorithmic_bytes"] == _counted_bytes(eng, ids, word)
    assert stats["elapsed_ms"] > 0 and stats["n_launches"] == 1


@pytest.mark.gpu
def test_gene_tree_scratch_regrows(require_gpu):
    """One engine: a small call, then more and longer elements (the kept
    scratch and length buffers grow), then the small call again; each against
    the oracle.  Releasing the kept scratch leaves the next call unchanged."""
    w = synth.make_gt_workload(seed=9, n_elements=24, S=24)
    opts = E.default_options(min_len=1)
    eng = E.Engine.from_workload(w, opts)
    ids, st = eng.add_elements_with_trees(w.align, w.starts, w.sizes, w.children())
    m = O.parse_mod(w.mod_text)[0]
    L = w.lengths()
    ref = np.array([O.gt_loglik(m, w.tree.S, g.children, g.seg_gene, g.seg_species, g.seg_len * w.scale, w.Z[c],
                                w.r1[c], w.r2[c], w.align, w.starts[c], w.sizes[c], gap_prop=opts.gap_prop)
                    for c, g in enumerate(w.genes)])
    order = np.argsort(w.sizes)
    small, big = order[:3], order
    for sel, prec, tol in ((small, E.PA_FP32, 1e-6), (big, E.PA_FP64, 1e-9), (small, E.PA_FP32, 1e-6)):
        np.testing.assert_allclose(eng.loglik_lengths(ids[sel], L[sel], prec), ref[sel], rtol=tol)
    eng.release_scratch()
    np.testing.assert_allclose(eng.loglik_lengths(ids[big], L[big], E.PA_FP32), ref[big], rtol=1e-6)
