"""Writes tests/golden/golden.npz: vectors that pin the oracle independently.

* expm: transition matrices exp(Q t) from scipy.linalg.expm for random GTR
  rate matrices and a range of t (including tiny and large t).
* brute: per-element log-likelihoods on small trees (4-5 tips) obtained by
  summing over every assignment of bases to the internal nodes, with the
  reference's leaf semantics (src/PhyloAcc-ST/bpp_c.hpp:213-261): a,c,g,t
  observed; r,y,k,m,s,w two-base sets; '-' and any other character
  integrate out (factor 1).  Branch rate 1 / r1 / r2 for Z = 0 / 1 / 2
  (bpp_c.cpp:661-663); branches of length <= 1e-10 use the identity
  (bpp.hpp:373-395).
* zcond: exact conditional distribution of the root's Z and of one
  child's Z given its parent, by enumeration over all Z configurations,
  for given per-branch log-emissions (the target of Update_Z_subtree,
  bpp_c.cpp:1182-1293).

Run:  python tests/golden/gen_golden.py   (numpy + scipy; deterministic seeds)
"""
import itertools
import os
import sys

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from phyloacc_amd import synth  # noqa: E402  (tree / GTR generators only)

AMB = {"r": (0, 2), "y": (1, 3), "k": (2, 3), "m": (0, 1), "s": (1, 2), "w": (0, 3)}
BASE = {"a": 0, "c": 1, "g": 2, "t": 3}


def trans(Q, t):
    if t <= 1e-10:
        return np.eye(4)
    return scipy.linalg.expm(Q * t)


def brute_loglik(tree, pi, Q, cols, Z, r1, r2):
    """cols: list of per-site strings (one char per leaf, leaf order)."""
    S, N = tree.S, tree.N
    rate = {0: 1.0, 1: r1, 2: r2}
    P = {s: trans(Q, tree.dist[s] * rate[int(Z[s])]) for s in range(N - 1)}
    internal = list(range(S, N))
    total = 0.0
    for col in cols:
        Lg = 0.0
        for x in itertools.product(range(4), repeat=len(internal)):
            st = dict(zip(internal, x))
            v = pi[st[N - 1]]
            for s in range(N - 1):
                p = tree.parent[s]
                if s >= S:
                    v *= P[s][st[p], st[s]]
                else:
                    ch = col[s]
                    if ch in BASE:
                        v *= P[s][st[p], BASE[ch]]
                    elif ch in AMB:
                        b1, b2 = AMB[ch]
                        v *= P[s][st[p], b1] + P[s][st[p], b2]
            Lg += v
        total += np.log(Lg)
    return total


def z_enumerate(parent, nodes, fixZ, E, g, l, l2):
    """Exact P(Z) over all configurations of the element's nodes (root last in nodes)."""
    root = nodes[-1]
    T = np.zeros((3, 3))       # T[i, j] = P(child i | parent j), free nodes
    T[:, 0] = [1 - g, g, 0]
    T[:, 1] = [0, 1 - l, l]
    T[:, 2] = [0, l2, 1 - l2]
    F = T.copy()               # fixZ nodes: no gains into 2, conserved stays conserved
    F[:, 1] = [0, 1, 0]
    F[:, 2] = [1, 1, 1]
    confs, ws = [], []
    for z in itertools.product(range(3), repeat=len(nodes)):
        Zd = dict(zip(nodes, z))
        if Zd[root] == 2:
            continue
        if any(fixZ[s] and Zd[s] == 2 for s in nodes):
            continue
        w = 0.5
        for s in nodes[:-1]:
            M = F if fixZ[s] else T
            w *= M[Zd[s], Zd[parent[s]]] * np.exp(E[s][Zd[s]])
        if w > 0:
            confs.append(z)
            ws.append(w)
    ws = np.array(ws) / np.sum(ws)
    return np.array(confs), ws


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    # expm goldens
    Qs, ts, Ps = [], [], []
    for k in range(4):
        pi, Q = synth.gtr(rng)
        for t in (1e-9, 1e-4, 0.01, 0.1, 0.5, 2.0, 10.0):
            Qs.append(Q)
            ts.append(t)
            Ps.append(scipy.linalg.expm(Q * t))
    out["expm_Q"] = np.array(Qs)
    out["expm_t"] = np.array(ts)
    out["expm_P"] = np.array(Ps)
    # brute-force likelihood cases
    cases = []
    chars = list("acgt" * 6 + "-n" + "rykmsw")
    for k in range(6):
        S = 4 + (k % 2)
        tree = synth.random_tree(S, 0.2, rng)
        pi, Q = synth.gtr(rng)
        G = 12
        cols = ["".join(rng.choice(chars) for _ in range(S)) for _ in range(G)]
        if k == 0:
            cols[0] = "-" * S       # an all-gap column contributes log 1
        Z = np.full(tree.N, 1, np.int32)
        for s in range(tree.N - 1):
            Z[s] = rng.integers(0, 3)
        Z[tree.N - 1] = rng.integers(0, 2)
        r1, r2 = float(rng.uniform(0.1, 0.9)), float(rng.uniform(1.2, 4.0))
        ll = brute_loglik(tree, pi, Q, cols, Z, r1, r2)
        cases.append(dict(mod=synth.mod_text(pi, Q, tree), rows=["".join(c[s] for c in cols) for s in range(S)],
                          Z=Z, r1=r1, r2=r2, loglik=ll, S=S))
    for i, c in enumerate(cases):
        out[f"brute{i}_mod"] = np.array(c["mod"])
        out[f"brute{i}_rows"] = np.array(c["rows"])
        out[f"brute{i}_Z"] = c["Z"]
        out[f"brute{i}_rates"] = np.array([c["r1"], c["r2"]])
        out[f"brute{i}_loglik"] = np.array(c["loglik"])
    out["n_brute"] = np.array(len(cases))
    # exact Z conditionals on a 4-tip tree (7 nodes)
    tree = synth.random_tree(4, 0.2, rng)
    nodes = []

    def post(v):
        if v >= tree.S:
            post(tree.children[v][0])
            post(tree.children[v][1])
        nodes.append(v)

    post(tree.N - 1)
    E = rng.normal(-20, 3, size=(tree.N, 3))
    for fz_case in range(2):
        fixZ = np.zeros(tree.N, np.int32)
        fixZ[tree.N - 1] = 1
        if fz_case == 1:
            fixZ[0] = 1
            fixZ[tree.parent[0]] = 1
        confs, ws = z_enumerate(tree.parent, nodes, fixZ, E, 0.4, 0.2, 0.1)
        out[f"zc{fz_case}_confs"] = confs
        out[f"zc{fz_case}_w"] = ws
        out[f"zc{fz_case}_fixZ"] = fixZ
    out["zc_nodes"] = np.array(nodes)
    out["zc_parent"] = tree.parent
    out["zc_E"] = E
    out["zc_mod"] = np.array(synth.mod_text(*synth.gtr(rng), tree))
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
