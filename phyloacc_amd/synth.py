"""Seeded synthetic PhyloAcc-ST workloads (trees, GTR models, alignments).

The benchmark configurations are described by their shapes only (species
count, branch-length law, element-length law, gap rate, rate classes), so the
inputs here are simulated: a random rooted binary tree, a GTR substitution
model, and per-element alignments evolved down the tree with per-branch rate
multipliers (neutral 1, conserved r1, accelerated r2).

Node numbering follows the reference's ``.mod`` reader
(src/PhyloAcc-ST/newick.cpp:382-428): post-order, leaves 0..S-1 in encounter
order, internal nodes S..N-1 in post-order, root N-1, children ordered by id
(src/PhyloAcc-ST/bpp.cpp:105-117).  ``Workload.mod_text`` is a ``.mod`` file
in the reference's format (BACKGROUND / RATE_MAT / TREE with every node
labelled, newick.cpp:416-419).
"""
from __future__ import annotations

import dataclasses
import numpy as np

CODES = np.frombuffer(b"acgt", dtype=np.uint8)


@dataclasses.dataclass
class Tree:
    S: int
    children: np.ndarray      # [N, 2] int32, -1 for leaves
    parent: np.ndarray        # [N] int32, -1 for root
    dist: np.ndarray          # [N] float64 (float32-rounded, as the reference parser stores it)
    names: list
    newick: str

    @property
    def N(self):
        return 2 * self.S - 1


def random_tree(S: int, mean_bl: float, rng: np.random.Generator, prefix="sp") -> Tree:
    """Yule-style random rooted binary tree with Exp(mean_bl) branch lengths."""
    # build as nested lists: node = [name or None, [kids], bl]
    root = {"kids": [], "bl": 0.0}
    leaves = [root]
    while len(leaves) < S:
        i = int(rng.integers(len(leaves)))
        v = leaves.pop(i)
        a = {"kids": [], "bl": 0.0}
        b = {"kids": [], "bl": 0.0}
        v["kids"] = [a, b]
        leaves += [a, b]
    for v in _walk(root):
        if v is not root:
            v["bl"] = float(np.float32(rng.exponential(mean_bl)))
    return _finish(root, S, prefix)


def _walk(v):
    yield v
    for k in v["kids"]:
        yield from _walk(k)


def _finish(root, S, prefix):
    # reference numbering: post-order, leaves and internal nodes counted apart
    leaf_id = [0]
    int_id = [0]

    def number(v):
        for k in v["kids"]:
            number(k)
        if not v["kids"]:
            v["id"] = leaf_id[0]
            leaf_id[0] += 1
        else:
            v["id"] = S + int_id[0]
            int_id[0] += 1

    number(root)
    N = 2 * S - 1
    children = -np.ones((N, 2), np.int32)
    parent = -np.ones(N, np.int32)
    dist = np.zeros(N, np.float64)
    names = [""] * N
    for v in _walk(root):
        i = v["id"]
        dist[i] = v["bl"]
        names[i] = f"{prefix}{i}" if not v["kids"] else f"n{i}"
        if v["kids"]:
            a, b = sorted(k["id"] for k in v["kids"])
            children[i] = (a, b)
            parent[a] = i
            parent[b] = i

    def nw(v):
        if not v["kids"]:
            return f"{names[v['id']]}:{v['bl']:.9g}"
        inner = ",".join(nw(k) for k in v["kids"])
        if v is root:
            return f"({inner}){names[v['id']]};"
        return f"({inner}){names[v['id']]}:{v['bl']:.9g}"

    return Tree(S, children, parent, dist, names, nw(root))


def gtr(rng: np.random.Generator):
    """pi ~ Dirichlet(10,10,10,10); exchangeabilities ~ LogNormal(0, 0.5);
    Q normalised to one expected substitution per unit time."""
    pi = rng.dirichlet([10.0] * 4)
    ex = rng.lognormal(0.0, 0.5, size=6)
    R = np.zeros((4, 4))
    k = 0
    for i in range(4):
        for j in range(i + 1, 4):
            R[i, j] = R[j, i] = ex[k]
            k += 1
    Q = R * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(1))
    Q /= -(pi * np.diag(Q)).sum()
    return pi, Q


def mod_text(pi, Q, tree: Tree) -> str:
    lines = ["ALPHABET: A C G T", "ORDER: 0", "SUBST_MOD: REV",
             "BACKGROUND: " + " ".join(f"{x:.17g}" for x in pi), "RATE_MAT:"]
    for i in range(4):
        lines.append("  " + " ".join(f"{x:.17g}" for x in Q[i]))
    lines.append("TREE: " + tree.newick)
    return "\n".join(lines) + "\n"


def _eig(pi, Q):
    d = np.sqrt(pi)
    A = (Q * d[:, None]) / d[None, :]
    A = 0.5 * (A + A.T)
    lam, U = np.linalg.eigh(A)
    V = U / d[:, None]          # right eigenvectors
    Vi = U.T * d[None, :]       # inverse
    return lam, V, Vi


def trans_batch(lam, V, Vi, t):
    """P(t) for a vector of times t -> [len(t), 4, 4] (row = from)."""
    t = np.asarray(t, np.float64)
    E = np.exp(lam[None, :] * t[:, None])
    P = (V[None, :, :] * E[:, None, :]) @ Vi
    return np.clip(P, 0.0, None)


@dataclasses.dataclass
class Workload:
    name: str
    tree: Tree
    pi: np.ndarray
    Q: np.ndarray
    mod_text: str
    align: np.ndarray         # [S, G] uint8 lower-case characters, rows in tree leaf order
    sizes: np.ndarray         # [C] element lengths (BED end - start)
    starts: np.ndarray        # [C] cumulative starts
    outgroup: list
    conserve: list
    target: list
    rates: np.ndarray         # [C, N] simulated rate multiplier per branch
    seed: int

    @property
    def C(self):
        return len(self.sizes)


def _lengths(kind, C, rng):
    if kind == "uniform100_500":
        return rng.integers(100, 501, size=C)
    if kind == "lognormal250":
        x = rng.lognormal(np.log(250.0), 0.6, size=C)
        return np.clip(np.rint(x), 50, 3000).astype(np.int64)
    if kind == "pareto":
        a, lo, hi = 1.5, 50.0, 10000.0
        u = rng.random(C)
        x = lo / (1.0 - u * (1.0 - (lo / hi) ** a)) ** (1.0 / a)
        return np.clip(np.floor(x), 50, 10000).astype(np.int64)
    if kind == "lognormal300":
        x = rng.lognormal(np.log(300.0), 0.6, size=C)
        return np.clip(np.rint(x), 50, 10000).astype(np.int64)
    raise ValueError(kind)


def _clade_tips(tree, v):
    out = []
    stack = [v]
    while stack:
        x = stack.pop()
        if x < tree.S:
            out.append(x)
        else:
            stack += list(tree.children[x])
    return sorted(out)


def simulate(tree, pi, Q, sizes, rates, gap_frac, rng):
    """Evolve every element down the tree.  rates: [C, N] per-branch multipliers."""
    lam, V, Vi = _eig(pi, Q)
    C = len(sizes)
    G = int(np.sum(sizes))
    elem = np.repeat(np.arange(C), sizes)
    elem16 = elem.astype(np.int64) * 16
    N = tree.N
    states = np.empty((N, G), np.uint8)
    cpi = np.cumsum(pi)
    states[N - 1] = np.searchsorted(cpi, rng.random(G) * cpi[-1], side="right").clip(0, 3)
    order = []
    stack = [N - 1]
    while stack:
        v = stack.pop()
        order.append(v)
        if v >= tree.S:
            stack += list(tree.children[v])
    for v in order:
        if v == N - 1:
            continue
        p = tree.parent[v]
        P = trans_batch(lam, V, Vi, tree.dist[v] * rates[:, v])        # [C,4,4]
        cum = np.cumsum(P, axis=2).reshape(-1)                          # flat [C*16]
        base = elem16 + states[p].astype(np.int64) * 4
        u = rng.random(G) * np.take(cum, base + 3)
        st = (u >= np.take(cum, base)).astype(np.uint8)
        st += u >= np.take(cum, base + 1)
        st += u >= np.take(cum, base + 2)
        states[v] = st
    align = CODES[states[: tree.S]]
    if gap_frac > 0:
        mask = rng.random(align.shape) < gap_frac
        gapc = np.where(rng.random(align.shape) < 0.5, ord("-"), ord("n")).astype(np.uint8)
        align = np.where(mask, gapc, align)
    return np.ascontiguousarray(align)


def make_workload(config: int, seed: int = 1, n_elements: int | None = None, lengths=None) -> Workload:
    """Build the synthetic inputs of one benchmark configuration.

    1: S=10, Exp(0.1) branches, 200 elements U[100,500], 2% gap/N, types
       neutral / conserved (r1=0.3) / accelerated (r2=2.0 on 2 target tips).
    2: S=43, Exp(0.08), 5000 elements LogNormal(ln 250, 0.6) in [50,3000],
       5% gaps, 4 target tips, same three types.
    3: S=43, 50000 elements Pareto(1.5) on [50, 10000]; 1-20% of branches
       accelerated per element, r1 ~ U[0.1,0.5], r2 ~ U[1.5,4]; 2% gaps.
    ``n_elements`` scales the element count down (tests) keeping the laws.
    ``lengths`` replaces the length draws by the given element lengths (tests
    that need chosen lengths under a configuration's other laws).
    """
    rng = np.random.default_rng(seed)
    if config == 1:
        S, mbl, C, lk, gap, ntarget = 10, 0.1, 200, "uniform100_500", 0.02, 2
    elif config in (2, 4):
        S, mbl, C, lk, gap, ntarget = 43, 0.08, 5000, "lognormal250", 0.05, 4
    elif config == 3:
        S, mbl, C, lk, gap, ntarget = 43, 0.08, 50000, "pareto", 0.02, 4
    else:
        raise ValueError(config)
    if n_elements is not None:
        C = int(n_elements)
    if lengths is not None:
        C = len(lengths)
    tree = random_tree(S, mbl, rng)
    pi, Q = gtr(rng)
    sizes = _lengths(lk, C, rng)
    if lengths is not None:
        sizes = np.asarray(lengths, np.int64).copy()
    N = tree.N
    # groups: the smaller root clade is the outgroup; targets are tips of the other side
    r = N - 1
    a, b = tree.children[r]
    ca, cb = _clade_tips(tree, a), _clade_tips(tree, b)
    small, big = (ca, cb) if len(ca) <= len(cb) else (cb, ca)
    outgroup = small[:2]
    target = [int(x) for x in rng.choice(big, size=min(ntarget, len(big) - 1), replace=False)]
    conserve = [x for x in range(S) if x not in target and x not in outgroup]
    rates = np.ones((C, N))
    if config in (1, 2, 4):
        typ = rng.integers(0, 3, size=C)       # 0 neutral, 1 conserved, 2 accelerated
        rates[typ == 1] = 0.3
        rates[typ == 2] = 0.3
        for t in target:
            rates[typ == 2, t] = 2.0
    else:
        r1 = rng.uniform(0.1, 0.5, size=C)
        r2 = rng.uniform(1.5, 4.0, size=C)
        rates[:] = r1[:, None]
        frac = rng.uniform(0.01, 0.20, size=C)
        nacc = np.maximum(1, np.rint(frac * (N - 1))).astype(int)
        # nacc[c] distinct random branches per element: ranks of a random key
        rank = np.argsort(np.argsort(rng.random((C, N - 1)), axis=1), axis=1)
        acc = rank < nacc[:, None]
        rates[:, : N - 1] = np.where(acc, r2[:, None], r1[:, None])
    align = simulate(tree, pi, Q, sizes, rates, gap, rng)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    return Workload(f"config{config}", tree, pi, Q, mod_text(pi, Q, tree), align,
                    sizes.astype(np.int64), starts, [int(x) for x in outgroup],
                    [int(x) for x in conserve], target, rates, seed)


def pareto_quantiles(k: int, alpha: float = 1.5, lo: float = 50.0, hi: float = 10000.0):
    """k lengths at the mid-quantiles (i + 0.5) / k of the truncated Pareto
    law of the 50,000-element configuration (inverse of its CDF)."""
    u = (np.arange(k) + 0.5) / k
    x = lo / (1.0 - u * (1.0 - (lo / hi) ** alpha)) ** (1.0 / alpha)
    return np.clip(np.floor(x), 50, 10000).astype(np.int64)


def small_workload(S=6, C=4, length=40, seed=3, gap=0.05, mean_bl=0.15):
    """Tiny workload for exhaustive checks (brute force, golden vectors)."""
    rng = np.random.default_rng(seed)
    tree = random_tree(S, mean_bl, rng)
    pi, Q = gtr(rng)
    sizes = np.full(C, length, np.int64)
    rates = np.ones((C, tree.N))
    align = simulate(tree, pi, Q, sizes, rates, gap, rng)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    a, b = tree.children[tree.N - 1]
    ca = _clade_tips(tree, a)
    cb = _clade_tips(tree, b)
    small_, big = (ca, cb) if len(ca) <= len(cb) else (cb, ca)
    return Workload("small", tree, pi, Q, mod_text(pi, Q, tree), align, sizes, starts,
                    small_[:1], [x for x in big[1:]], big[:1], rates, seed)


# ----------------------------------------------------------------- gene trees

@dataclasses.dataclass
class GeneTree:
    """One element's gene tree embedded in the species tree (PhyloAcc-GT style).

    Leaves 0..S-1 are the species' sequences; internal gene nodes S..N-1 are
    numbered in post-order of the gene tree; root N-1.  seg_* list, for every
    gene branch, the pieces it spends in each species branch: (gene child,
    species branch, length in coalescent units)."""
    children: np.ndarray      # [N, 2]
    parent: np.ndarray        # [N]
    height: np.ndarray        # [N] coalescent-unit heights
    seg_gene: np.ndarray      # [K] gene node whose parent branch holds the piece
    seg_species: np.ndarray   # [K] species branch (node id of its lower end)
    seg_len: np.ndarray       # [K] length in coalescent units


def ultrametric_tree(S, mean_interval, rng, prefix="sp", mean_bl=None):
    """Random ultrametric species tree: lineages merge pairwise at Exp(mean)
    intervals; with ``mean_bl`` all heights are rescaled to that mean branch length."""
    lineages = [{"kids": [], "h": 0.0} for _ in range(S)]
    h = 0.0
    while len(lineages) > 1:
        h += rng.exponential(mean_interval)
        i, j = rng.choice(len(lineages), 2, replace=False)
        a, b = lineages[i], lineages[j]
        v = {"kids": [a, b], "h": h}
        lineages = [x for k, x in enumerate(lineages) if k not in (i, j)] + [v]
    root = lineages[0]
    if mean_bl is not None:
        tot = sum(v["h"] - k["h"] for v in _walk(root) for k in v["kids"])
        f = mean_bl * (2 * S - 2) / tot
        for v in _walk(root):
            v["h"] *= f
    for v in _walk(root):
        for k in v["kids"]:
            k["bl"] = float(np.float32(v["h"] - k["h"]))
    root["bl"] = 0.0
    t = _finish(root, S, prefix)
    heights = np.zeros(t.N)
    for v in _walk(root):
        heights[v["id"]] = v["h"]
    return t, heights


def msc_gene_tree(tree: Tree, sp_height, rng) -> GeneTree:
    """Multispecies-coalescent draw: within species branch s, k lineages
    coalesce at rate k(k-1)/2 per coalescent unit until the branch's top."""
    S, N = tree.S, tree.N
    post = []
    stack = [(N - 1, 0)]
    while stack:
        v, k = stack.pop()
        if v >= S and k == 0:
            stack.append((v, 1))
            stack.append((tree.children[v][1], 0))
            stack.append((tree.children[v][0], 0))
        else:
            post.append(v)
    # gene nodes: dicts with kids, height, species branch where they were born
    nodes = [{"kids": [], "h": 0.0, "sp": s, "id": s} for s in range(S)]
    pieces = []                      # (gene node index, species branch, length)
    lineages = {s: [s] for s in range(S)}   # species node -> gene node indices entering its branch
    for s in post:
        if s >= S:
            a, b = tree.children[s]
            lin = lineages[a] + lineages[b]
        else:
            lin = list(lineages[s])
        bottom = sp_height[s]
        top = sp_height[tree.parent[s]] if s != N - 1 else np.inf
        h = bottom
        enter = {g: bottom for g in lin}
        while len(lin) > 1:
            k = len(lin)
            h_next = h + rng.exponential(2.0 / (k * (k - 1)))
            if h_next >= top:
                break
            i, j = rng.choice(k, 2, replace=False)
            gi, gj = lin[i], lin[j]
            nid = len(nodes)
            nodes.append({"kids": [gi, gj], "h": h_next, "sp": s})
            for g in (gi, gj):
                pieces.append((g, s, h_next - max(enter[g], nodes[g]["h"])))
            lin = [x for q, x in enumerate(lin) if q not in (i, j)] + [nid]
            enter[nid] = h_next
            h = h_next
        if s != N - 1:
            for g in lin:
                pieces.append((g, s, top - max(enter[g], nodes[g]["h"])))
            lineages[s] = lin
    # renumber internal gene nodes in gene-tree post-order (leaves keep species ids)
    groot = len(nodes) - 1
    order = []
    st = [(groot, 0)]
    while st:
        v, k = st.pop()
        if nodes[v]["kids"] and k == 0:
            st.append((v, 1))
            st.append((nodes[v]["kids"][1], 0))
            st.append((nodes[v]["kids"][0], 0))
        else:
            order.append(v)
    newid = {}
    nxt = S
    for v in order:
        if nodes[v]["kids"]:
            newid[v] = nxt
            nxt += 1
        else:
            newid[v] = v
    children = -np.ones((N, 2), np.int32)
    parent = -np.ones(N, np.int32)
    height = np.zeros(N)
    for v, d in enumerate(nodes):
        i = newid[v]
        height[i] = d["h"]
        if d["kids"]:
            a, b = sorted(newid[k] for k in d["kids"])
            children[i] = (a, b)
            parent[a] = i
            parent[b] = i
    seg_gene = np.array([newid[g] for g, _, _ in pieces], np.int32)
    seg_sp = np.array([s for _, s, _ in pieces], np.int32)
    seg_len = np.array([l for _, _, l in pieces], np.float64)
    return GeneTree(children, parent, height, seg_gene, seg_sp, seg_len)


def gene_effective_lengths(gt: GeneTree, Z, r1, r2, scale):
    """Substitution length of every gene branch: sum of its pieces' lengths
    times the rate class of the species branch holding each piece."""
    rate = np.where(Z[gt.seg_species] == 2, r2, np.where(Z[gt.seg_species] == 1, r1, 1.0))
    N = len(gt.parent)
    eff = np.zeros(N)
    np.add.at(eff, gt.seg_gene, gt.seg_len * rate * scale)
    return eff


@dataclasses.dataclass
class GTWorkload:
    """Elements that each carry their own gene tree (PhyloAcc-GT style)."""
    name: str
    tree: Tree                 # species tree, lengths in coalescent units
    sp_height: np.ndarray      # [N] species node heights
    pi: np.ndarray
    Q: np.ndarray
    mod_text: str
    align: np.ndarray          # [S, G]
    sizes: np.ndarray          # [C]
    starts: np.ndarray         # [C]
    genes: list                # [C] GeneTree
    Z: np.ndarray              # [C, N] simulated species-branch classes
    r1: np.ndarray             # [C]
    r2: np.ndarray             # [C]
    scale: float               # substitutions per coalescent unit
    seed: int

    outgroup = ()              # no outgroup collapse and no CONSERVE filter for gene trees
    conserve = ()

    @property
    def C(self):
        return len(self.sizes)

    def children(self):
        return np.stack([g.children for g in self.genes]).astype(np.int32)

    def lengths(self, Z=None, r1=None, r2=None):
        """[C, N] substitution lengths of every gene branch."""
        Z = self.Z if Z is None else Z
        r1 = self.r1 if r1 is None else r1
        r2 = self.r2 if r2 is None else r2
        return np.stack([gene_effective_lengths(g, Z[c], r1[c], r2[c], self.scale)
                         for c, g in enumerate(self.genes)])


def simulate_gene_trees(children, lengths, pi, Q, sizes, gap_frac, rng, S):
    """Evolve every element down its own tree.  Ids are post-order (a parent's
    id exceeds its children's), so visiting ids from the root down is a
    valid top-down order for every element at once."""
    lam, V, Vi = _eig(pi, Q)
    C, N = lengths.shape
    G = int(np.sum(sizes))
    elem = np.repeat(np.arange(C), sizes)
    elem16 = elem.astype(np.int64) * 16
    parent = np.full((C, N), -1, np.int64)
    for q in range(2):
        np.put_along_axis(parent, children[:, S:, q].astype(np.int64), np.arange(S, N)[None, :].repeat(C, 0), 1)
    states = np.empty((N, G), np.uint8)
    flat = states.reshape(-1)
    cpi = np.cumsum(pi)
    states[N - 1] = np.searchsorted(cpi, rng.random(G) * cpi[-1], side="right").clip(0, 3)
    cols = np.arange(G, dtype=np.int64)
    for v in range(N - 2, -1, -1):
        P = trans_batch(lam, V, Vi, lengths[:, v])
        cum = np.cumsum(P, axis=2).reshape(-1)
        # each element's sites read one parent row: contiguous runs of the gather
        par = np.take(flat, np.repeat(parent[:, v] * G, sizes) + cols)
        base = elem16 + par.astype(np.int64) * 4
        u = rng.random(G) * np.take(cum, base + 3)
        st = (u >= np.take(cum, base)).astype(np.uint8)
        st += u >= np.take(cum, base + 1)
        st += u >= np.take(cum, base + 2)
        states[v] = st
    align = CODES[states[:S]]
    if gap_frac > 0:
        mask = rng.random(align.shape) < gap_frac
        gapc = np.where(rng.random(align.shape) < 0.5, ord("-"), ord("n")).astype(np.uint8)
        align = np.where(mask, gapc, align)
    return np.ascontiguousarray(align)


def _gt_genes(args):
    tree, h, seed, lo, hi = args
    return [msc_gene_tree(tree, h, np.random.default_rng([seed, 1, c])) for c in range(lo, hi)]


def _gt_align(args):
    children, lengths, pi, Q, sizes, gap, seed, k, S = args
    return simulate_gene_trees(children, lengths, pi, Q, sizes, gap, np.random.default_rng([seed, 2, k]), S)


GT_CHUNK = 500     # elements per independently seeded simulation chunk


def make_gt_workload(seed: int = 5, n_elements: int = 20000, S: int = 120, mean_bl: float = 0.5,
                     scale: float = 0.1, gap: float = 0.02, workers: int = 1) -> GTWorkload:
    """Gene-tree workload: an ultrametric species tree in coalescent units whose
    branch lengths average ``mean_bl``; every element draws its own gene tree
    from the multispecies coalescent; lengths LogNormal(ln 300, 0.6) clipped to
    [50, 3000]; 1-20% of species branches accelerated (r2 ~ U[1.5, 4]), the
    rest conserved (r1 ~ U[0.1, 0.5]); ``scale`` substitutions per unit.

    Element c's gene tree comes from its own stream (seed, 1, c) and the
    sequences of chunk k (elements k*GT_CHUNK ...) from stream (seed, 2, k),
    so the workload is the same for any number of ``workers`` (forked
    processes; use them only before the process has touched the GPU)."""
    rng = np.random.default_rng(seed)
    tree, h = ultrametric_tree(S, 1.0, rng, mean_bl=mean_bl)
    N = tree.N
    pi, Q = gtr(rng)
    C = int(n_elements)
    sizes = np.clip(np.rint(rng.lognormal(np.log(300), 0.6, C)), 50, 3000).astype(np.int64)
    r1 = rng.uniform(0.1, 0.5, size=C)
    r2 = rng.uniform(1.5, 4.0, size=C)
    frac = rng.uniform(0.01, 0.20, size=C)
    nacc = np.maximum(1, np.rint(frac * (N - 1))).astype(int)
    rank = np.argsort(np.argsort(rng.random((C, N - 1)), axis=1), axis=1)
    Z = np.ones((C, N), np.int32)
    Z[:, : N - 1] = np.where(rank < nacc[:, None], 2, 1)
    bounds = [(lo, min(C, lo + GT_CHUNK)) for lo in range(0, C, GT_CHUNK)]
    pool = None
    if workers > 1 and len(bounds) > 1:
        import multiprocessing as mp
        pool = mp.get_context("fork").Pool(min(workers, len(bounds)))
    try:
        run = pool.map if pool else (lambda f, xs: [f(x) for x in xs])
        genes = [g for part in run(_gt_genes, [(tree, h, seed, lo, hi) for lo, hi in bounds]) for g in part]
        w = GTWorkload(f"gt{S}", tree, h, pi, Q, mod_text(pi, Q, tree), None, sizes,
                       np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64), genes, Z, r1, r2, scale, seed)
        ch, L = w.children(), w.lengths()
        parts = run(_gt_align, [(ch[lo:hi], L[lo:hi], pi, Q, sizes[lo:hi], gap, seed, k, S)
                                for k, (lo, hi) in enumerate(bounds)])
    finally:
        if pool:
            pool.close()
            pool.join()
    w.align = np.ascontiguousarray(np.concatenate(parts, axis=1))
    return w
