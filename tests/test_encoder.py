"""Element encoding and tree indexing: the engine's host encoder (through the
C ABI, no GPU needed) against hand-derived known answers and against the
oracle's independent restatement of the BPP_C constructor
(src/PhyloAcc-ST/bpp_c.hpp:143-476)."""
import re

import numpy as np
import pytest

from oracle import oracle_py as O
from phyloacc_amd import engine as E
from phyloacc_amd import synth

# ((A,B)n1,(C,D)n2)root -> ids A0 B1 C2 D3 n1=4 n2=5 root=6 (newick.cpp:382-400)
MOD4 = """BACKGROUND: 0.25 0.25 0.25 0.25
RATE_MAT:
 -1 0.333333333333333 0.333333333333333 0.333333333333334
 0.333333333333333 -1 0.333333333333333 0.333333333333334
 0.333333333333333 0.333333333333333 -1 0.333333333333334
 0.333333333333333 0.333333333333333 0.333333333333334 -1
TREE: ((A:0.1,B:0.2)n1:0.3,(C:0.4,D:0.5)n2:0.6)root;
"""


def f32(x):
    return float(np.float32(x))


def _engine(mod, outgroup, conserve, **kw):
    t = E.parse_mod(mod)
    opts = E.default_options(**kw)
    return E.Engine(t["S"], t["children"], t["parent"], t["dist"], t["pi"], t["Q"], outgroup, conserve, opts), t


def _align(rows):
    return np.array([np.frombuffer(r.encode(), np.uint8) for r in rows])


def test_library_exports_every_header_symbol():
    hdr = open(E.os.path.join(E.HERE, "..", "include", "phyloacc_gpu.h")).read()
    names = set(re.findall(r"\b(pa_[a-z_]+)\s*\(", hdr))
    L = E.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(E.EXPORTS) <= names


def test_mod_numbering_known_answer():
    t = E.parse_mod(MOD4)
    assert t["S"] == 4 and t["names"] == ["A", "B", "C", "D", "n1", "n2", "root"]
    np.testing.assert_array_equal(t["children"], [[-1, -1]] * 4 + [[0, 1], [2, 3], [4, 5]])
    np.testing.assert_array_equal(t["parent"], [4, 4, 5, 5, 6, 6, -1])
    np.testing.assert_array_equal(t["dist"], [f32(x) for x in (0.1, 0.2, 0.4, 0.5, 0.3, 0.6, 0.0)])
    # children are ordered by id, not by newick order (bpp.cpp:105-117)
    t2 = E.parse_mod(MOD4.replace("((A:0.1,B:0.2)n1:0.3,(C:0.4,D:0.5)n2:0.6)root", "((C:0.4,D:0.5)n2:0.6,A:0.7)root"))
    # leaves C0 D1 A2, n2=3, root=4 -> children of root sorted (2, 3)
    np.testing.assert_array_equal(t2["children"][4], [2, 3])
    assert t2["names"][:3] == ["C", "D", "A"]


def test_mod_parser_matches_oracle_and_generator():
    for cfg, n in ((1, 2), (2, 2)):
        w = synth.make_workload(cfg, seed=3, n_elements=n)
        t = E.parse_mod(w.mod_text)
        m, names = O.parse_mod(w.mod_text)
        np.testing.assert_array_equal(t["children"], w.tree.children)
        np.testing.assert_array_equal(t["dist"], np.array(m.dist[: m.N]))
        assert t["names"] == names == w.tree.names
        np.testing.assert_allclose(t["pi"], w.pi, rtol=1e-15)
        np.testing.assert_allclose(t["Q"], w.Q, rtol=1e-15)


def test_outgroup_collapse_known_answer():
    # A (outgroup) all gaps -> missing; n1 collapses onto B with length 0.2+0.3 (bpp_c.cpp:64-113)
    eng, t = _engine(MOD4, [0], [1, 2, 3], min_len=1)
    rows = ["-" * 8, "acgtacgt", "acgtaagt", "ccgtacgt"]
    ids, st = eng.add_elements(_align(rows), [0], [8])
    d = eng.element(ids[0])
    assert st[0] == E.PA_ELEM_OK and d["root"] == 6
    np.testing.assert_array_equal(d["missing"], [1, 0, 0, 0, 0, 0, 0])
    np.testing.assert_array_equal(d["nodes"], [1, 2, 3, 5, 6])
    np.testing.assert_array_equal(d["children"][6], [1, 5])
    assert d["parent"][1] == 6 and d["parent"][0] == -1 and d["parent"][6] == 7
    assert d["dist"][1] == f32(0.2) + f32(0.3)


def test_root_moves_when_outgroup_clade_missing():
    # A and B missing -> n1 missing; the root (in the outgroup lineage) moves to n2
    eng, t = _engine(MOD4, [0], [2, 3], min_len=1)
    rows = ["-" * 8, "-" * 8, "acgtaagt", "ccgtacgt"]
    ids, st = eng.add_elements(_align(rows), [0], [8])
    d = eng.element(ids[0])
    np.testing.assert_array_equal(d["missing"], [1, 1, 0, 0, 1, 0, 0])
    assert d["root"] == 5
    np.testing.assert_array_equal(d["nodes"], [2, 3, 5])
    assert d["dist"][5] == f32(0.6) + f32(0.0)


def test_prune_tree_collapses_everywhere():
    eng, t = _engine(MOD4, [0], [1, 3], min_len=1, prune_tree=1)
    rows = ["acgtacgt", "acgtacgt", "-" * 8, "ccgtacgt"]
    ids, st = eng.add_elements(_align(rows), [0], [8])
    d = eng.element(ids[0])
    np.testing.assert_array_equal(d["nodes"], [0, 1, 4, 3, 6])
    assert d["dist"][3] == f32(0.5) + f32(0.6)
    # without PRUNE_TREE the non-outgroup clade keeps its missing leaf
    eng2, _ = _engine(MOD4, [0], [1, 3], min_len=1, prune_tree=0)
    ids2, _ = eng2.add_elements(_align(rows), [0], [8])
    np.testing.assert_array_equal(eng2.element(ids2[0])["nodes"], [0, 1, 4, 2, 3, 5, 6])


def test_filters_min_len_conserve_and_trim():
    rows = ["acgtacgt", "acgt----", "--------", "acgtac--"]
    # MIN_LEN
    eng, _ = _engine(MOD4, [0], [1, 2, 3], min_len=9)
    _, st = eng.add_elements(_align(rows), [0], [8])
    assert st[0] == E.PA_ELEM_FILTERED
    # CONSERVE: C missing (8 gaps > 0.8*8); 1 of 3 > 0.3*3 -> filtered
    eng, _ = _engine(MOD4, [0], [1, 2, 3], min_len=1, conserve_prop=0.3)
    _, st = eng.add_elements(_align(rows), [0], [8])
    assert st[0] == E.PA_ELEM_FILTERED
    eng, _ = _engine(MOD4, [0], [1, 2, 3], min_len=1, conserve_prop=0.4)
    _, st = eng.add_elements(_align(rows), [0], [8])
    assert st[0] == E.PA_ELEM_OK
    # trimming: columns with more than 4*0.5 gap/N characters are dropped
    eng, _ = _engine(MOD4, [0], [1, 2, 3], min_len=1, trim_gap_percent=0.5)
    ids, st = eng.add_elements(_align(rows), [0], [8])
    d = eng.element(ids[0])
    assert d["GG"] == 6
    # gap counts use '-' only; 'n' does not make a species missing (bpp_c.hpp:304-314)
    rows_n = ["acgtacgt", "nnnnnnnn", "acgtacgt", "acgtacgt"]
    eng, _ = _engine(MOD4, [0], [1, 2, 3], min_len=1)
    ids, _ = eng.add_elements(_align(rows_n), [0], [8])
    assert eng.element(ids[0])["missing"][1] == 0


def test_codes_and_patterns():
    eng, _ = _engine(MOD4, [0], [1, 2, 3], min_len=1)
    rows = ["ACgtrykmsw-n?", "acgtacgtacgta", "acgtacgtacgta", "acgtacgtacgta"]
    ids, _ = eng.add_elements(_align(rows), [0], [13])
    d = eng.element(ids[0])
    np.testing.assert_array_equal(d["codes"][:, 0], [0, 1, 2, 3, 6, 7, 8, 9, 10, 11, 4, 5, 5])
    # patterns x weights reproduce the site columns exactly
    eng2, _ = _engine(MOD4, [0], [1, 2, 3], min_len=1)
    rows2 = ["aaaacccc", "aaaacccg", "aaaacccg", "aaaacccg"]
    ids2, _ = eng2.add_elements(_align(rows2), [0], [8])
    d2 = eng2.element(ids2[0])
    assert d2["P"] == 3
    np.testing.assert_array_equal(d2["site_pattern"], [0, 0, 0, 0, 1, 1, 1, 2])
    np.testing.assert_array_equal(d2["weights"], [4, 3, 1])


@pytest.mark.parametrize("kw", [dict(), dict(trim_gap_percent=0.2, gap_prop=0.1), dict(prune_tree=1, gap_prop=0.05),
                                dict(conserve_prop=0.05, gap_prop=0.1)])
def test_encoder_matches_oracle(kw):
    w = synth.make_workload(2, seed=9, n_elements=30)
    # plant heavy gaps in a few species so that missing / collapse / filters trigger
    rng = np.random.default_rng(2)
    al = w.align.copy()
    for s in rng.choice(w.tree.S, 6, replace=False):
        mask = rng.random(al.shape[1]) < 0.5
        al[s, mask] = ord("-")
    for s in w.outgroup:
        al[s, : al.shape[1] // 2] = ord("-")
    eng = E.Engine.from_workload(w, E.default_options(**kw))
    ids, st = eng.add_elements(al, w.starts, w.sizes)
    m = O.model_from_workload(w)
    oo = O.default_options(**kw)
    nf = 0
    for i in range(len(ids)):
        d = eng.element(ids[i])
        oe = O.Element(m, oo, al, w.starts[i], w.sizes[i])
        assert bool(d["filtered"]) == oe.filtered
        if oe.filtered:
            nf += 1
            continue
        a = oe.arrays()
        assert a["GG"] == d["GG"] and a["root"] == d["root"]
        np.testing.assert_array_equal(a["children"], d["children"])
        np.testing.assert_array_equal(a["parent"], d["parent"])
        np.testing.assert_array_equal(a["dist"], d["dist"])
        np.testing.assert_array_equal(a["missing"], d["missing"])
        np.testing.assert_array_equal(a["nodes"], d["nodes"])
        np.testing.assert_array_equal(a["codes"], d["codes"])
        assert d["weights"].sum() == d["GG"]


def test_rng_matches_oracle():
    for args in ((5, 0, 0, 0, 0, 0), (5, 3, 17, 999, (2 << 24) | (1 << 16), 2), (2**40 + 7, 1, 2, 3, 4, 3)):
        assert E.rng_uniform(*args) == O.lib().or_uniform(*args)
