"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It loads oracle/libpa_oracle.so, built by oracle/Makefile.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpa_oracle.so")
MAXN = 512


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class Options(C.Structure):
    _fields_ = [
        ("burnin", C.c_int), ("mcmc", C.c_int), ("adapt_freq", C.c_int),
        ("init_crate", C.c_double), ("init_nrate", C.c_double),
        ("cprior_a", C.c_double), ("cprior_b", C.c_double),
        ("nprior_a", C.c_double), ("nprior_b", C.c_double),
        ("ropt", C.c_int), ("cub", C.c_double), ("nlb", C.c_double),
        ("init_grate", C.c_double), ("init_lrate", C.c_double), ("init_lrate2", C.c_double),
        ("hyper_grate_a", C.c_double), ("hyper_grate_b", C.c_double),
        ("hyper_lrate_a", C.c_double), ("hyper_lrate_b", C.c_double),
        ("hyper_lrate2_a", C.c_double), ("hyper_lrate2_b", C.c_double),
        ("gap_prop", C.c_double), ("conserve_prop", C.c_double),
        ("constomis", C.c_double), ("nconstomis", C.c_double),
        ("trim_gap_percent", C.c_double), ("min_len", C.c_int), ("prune_tree", C.c_int),
        ("gapchar", C.c_int), ("seed", C.c_uint64), ("r1_shared_uniform", C.c_int),
    ]


def default_options(**kw) -> Options:
    """Reference defaults (src/PhyloAcc-ST/main.cpp:35-90)."""
    o = Options(burnin=200, mcmc=800, adapt_freq=500, init_crate=0.5, init_nrate=1.0,
                cprior_a=5, cprior_b=0.04, nprior_a=10, nprior_b=0.2, ropt=1, cub=1, nlb=0.6,
                init_grate=0.5, init_lrate=0.3, init_lrate2=0.1,
                hyper_grate_a=3, hyper_grate_b=1, hyper_lrate_a=1, hyper_lrate_b=9,
                hyper_lrate2_a=1, hyper_lrate2_b=9, gap_prop=0.8, conserve_prop=0.8,
                constomis=0.01, nconstomis=1.0, trim_gap_percent=1.0, min_len=50,
                prune_tree=0, gapchar=ord("-"), seed=5, r1_shared_uniform=1)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Model(C.Structure):
    _fields_ = [
        ("S", C.c_int), ("N", C.c_int),
        ("children", (C.c_int * 2) * MAXN), ("parent", C.c_int * MAXN),
        ("dist", C.c_double * MAXN), ("pi", C.c_double * 4), ("Q", C.c_double * 16),
        ("upper", C.c_int * MAXN), ("upper_conserve", C.c_int * MAXN),
        ("conserve", C.c_int * MAXN), ("n_conserve", C.c_int),
    ]


class ChainOut(C.Structure):
    _fields_ = [
        ("loglik", C.POINTER(C.c_double)), ("full", C.POINTER(C.c_double)),
        ("nrate", C.POINTER(C.c_double)), ("crate", C.POINTER(C.c_double)),
        ("grate", C.POINTER(C.c_double)), ("lrate", C.POINTER(C.c_double)),
        ("l2rate", C.POINTER(C.c_double)), ("Z", C.POINTER(C.c_byte)),
        ("postZ", C.POINTER(C.c_double)),
        ("max_loglik", C.c_double), ("max_m", C.c_int), ("max_Z", C.c_int * MAXN),
        ("vb", C.c_double), ("med", C.c_double * 5), ("failure", C.c_int),
        ("accept_n", C.c_int), ("accept_c", C.c_int), ("flags", C.c_int),
    ]


_lib = None


def gbgc_rate(Q0, s, b):
    """Q(s, b) of the gBGC model from Q0 [4, 4]."""
    Q0 = np.ascontiguousarray(Q0, np.float64).reshape(16)
    Q = np.zeros(16)
    lib().or_gbgc_rate(_ptr(Q0, C.c_double), float(s), float(b), _ptr(Q, C.c_double))
    return Q.reshape(4, 4)


class SpTree(C.Structure):
    _fields_ = [("S", C.c_int), ("N", C.c_int), ("parent", C.POINTER(C.c_int)),
                ("height", C.POINTER(C.c_double)), ("theta", C.POINTER(C.c_double))]


def sptree(S, parent, height, theta):
    """An or_sptree over numpy copies that live as long as the struct."""
    keep = [np.ascontiguousarray(parent, np.int32), np.ascontiguousarray(height, np.float64),
            np.ascontiguousarray(theta, np.float64)]
    t = SpTree(int(S), len(keep[0]), _ptr(keep[0], C.c_int), _ptr(keep[1], C.c_double), _ptr(keep[2], C.c_double))
    t._keep = keep
    return t


class GeneTree:
    """A gene tree as topology + heights (GTree): children [N, 2] (-1 at
    leaves), parent [N] (N at the root), heights [N], missing [N]."""

    def __init__(self, children, parent, heights, missing=None):
        self.ch = np.array(children, np.int32).reshape(-1)
        self.par = np.ascontiguousarray(parent, np.int32).copy()
        self.h = np.ascontiguousarray(heights, np.float64).copy()
        N = len(self.par)
        self.miss = np.zeros(N, np.int32) if missing is None else np.ascontiguousarray(missing, np.int32).copy()

    @property
    def root(self):
        return int(np.flatnonzero(self.par == len(self.par))[0])

    def args(self):
        return _ptr(self.ch, C.c_int), _ptr(self.par, C.c_int), _ptr(self.h, C.c_double)


def gt_species(sp, gt, v):
    return lib().or_gt_species(C.byref(sp), _ptr(gt.ch, C.c_int), _ptr(gt.h, C.c_double), int(v))


def gt_prior(sp, gt):
    return lib().or_gt_prior(C.byref(sp), *gt.args(), gt.root)


def gt_logpz(sp, gt, Z, c2m, n2m):
    z = np.ascontiguousarray(Z, np.int32)
    return lib().or_gt_logpz(C.byref(sp), _ptr(gt.ch, C.c_int), _ptr(gt.h, C.c_double), _ptr(gt.miss, C.c_int),
                             _ptr(z, C.c_int), float(c2m), float(n2m))


def gt_pieces(sp, gt):
    N = len(gt.par)
    cap = N * (N + 1)
    g = np.zeros(cap, np.int32)
    s = np.zeros(cap, np.int32)
    ln = np.zeros(cap)
    K = lib().or_gt_pieces(C.byref(sp), *gt.args(), gt.root, _ptr(g, C.c_int), _ptr(s, C.c_int), _ptr(ln, C.c_double))
    return g[:K], s[:K], ln[:K]


def _rows(align):
    keep = [np.ascontiguousarray(align[s]).tobytes() for s in range(align.shape[0])]
    return (C.c_char_p * len(keep))(*keep), keep


def gt_site_logliks(model, sp, gt, Z, r1, r2, align, start, length, piA=0.0, gapchar="-", gap_prop=0.8):
    """Per-site log-likelihoods of the gene tree at its heights (pieces at the
    species branches' rates), over sites [start, start + length)."""
    g, s, ln = gt_pieces(sp, gt)
    rows, keep = _rows(align)
    z = np.ascontiguousarray(Z, np.int32)
    out = np.zeros(int(length))
    m = model
    if piA > 0:
        raise ValueError("per-element pi: use gt_branch_len")
    lib().or_gt_site_logliks(C.byref(m), int(sp.S), _ptr(gt.ch, C.c_int), gt.root, len(g), _ptr(g, C.c_int),
                             _ptr(s, C.c_int), _ptr(ln, C.c_double), _ptr(z, C.c_int), float(r1), float(r2), rows,
                             int(start), int(length), ord(gapchar), float(gap_prop), _ptr(out, C.c_double))
    return out


def gt_branch_len(model, sp, gt, Z, c_rate, n_rate, c2m, n2m, gnode, delta, indicator, lens, loglik, logpz, prior,
                  seed, m, slot, align, start, length, piA=0.0, gapchar="-", gap_prop=0.8):
    """One Sample_BranchLen move; gt.h, loglik [2], logpz [2] are updated in
    place.  Returns (code, prior): code 1 accepted, 0 not, -1 error."""
    rows, keep = _rows(align)
    z = np.ascontiguousarray(Z, np.int32)
    ln = np.ascontiguousarray(lens, np.int32)
    assert loglik.dtype == np.float64 and logpz.dtype == np.float64
    pr = C.c_double(prior)
    code = lib().or_gt_branch_len(C.byref(model), float(piA), C.byref(sp), *gt.args(), _ptr(gt.miss, C.c_int),
                                  gt.root, _ptr(z, C.c_int), float(c_rate), float(n_rate), float(c2m), float(n2m),
                                  int(gnode), float(delta), int(indicator), _ptr(ln, C.c_int),
                                  _ptr(loglik, C.c_double), _ptr(logpz, C.c_double), C.byref(pr), int(seed), int(m),
                                  int(slot), rows, int(start), int(length), ord(gapchar), float(gap_prop))
    return code, pr.value


def gt_sample_history(model, sp, gt, Z, c_rate, n_rate, seed, m, length, align, start, size, piA=0.0,
                      gapchar="-", gap_prop=0.8):
    """Update_Tg(tosample=true): int8 [N, length] history of the first
    `length` sites of an element of `size` sites."""
    rows, keep = _rows(align)
    z = np.ascontiguousarray(Z, np.int32)
    out = np.zeros((len(gt.par), int(length)), np.int8)
    lib().or_gt_sample_history(C.byref(model), float(piA), C.byref(sp), *gt.args(), gt.root, _ptr(z, C.c_int),
                               float(c_rate), float(n_rate), int(seed), int(m), int(length), rows, int(start),
                               int(size), ord(gapchar), float(gap_prop), _ptr(out, C.c_int8))
    return out


def gt_sample_tree(model, sp, gt, Z, c_rate, n_rate, c2m, n2m, branch, indicator, lens, Tg, loglik, logpz, seed, m,
                   slot, align, start, size, root_ingrp=None, piA=0.0, gapchar="-", gap_prop=0.8, exact=False,
                   scores=None):
    """One Sample_tree2 move on gt (updated in place when accepted); Tg is
    the history of the first lens[indicator] sites.  Returns 1 or 0.  exact:
    the test hook's complete Hastings ratio; scores: a list that receives
    [(candidate, score), ...]."""
    rows, keep = _rows(align)
    N = len(gt.par)
    cn = np.zeros(N + 2, np.int32)
    cs = np.zeros(N + 2)
    nc = C.c_int(0)
    z = np.ascontiguousarray(Z, np.int32)
    ln = np.ascontiguousarray(lens, np.int32)
    T = np.ascontiguousarray(Tg, np.int8)
    assert T.shape == (len(gt.par), int(ln[indicator]))
    root = C.c_int(gt.root)
    ri = sp.N - 1 if root_ingrp is None else int(root_ingrp)
    code = lib().or_test_sample_tree(C.byref(model), float(piA), C.byref(sp), *gt.args(), C.byref(root),
                                     _ptr(z, C.c_int), float(c_rate), float(n_rate), float(c2m), float(n2m), ri,
                                     int(branch), int(indicator), _ptr(ln, C.c_int), _ptr(T, C.c_int8),
                                     _ptr(loglik, C.c_double), _ptr(logpz, C.c_double), int(seed), int(m), int(slot),
                                     rows, int(start), int(size), ord(gapchar), float(gap_prop), int(bool(exact)),
                                     _ptr(cn, C.c_int), _ptr(cs, C.c_double), C.byref(nc))
    if scores is not None:
        scores[:] = list(zip(cn[:nc.value].tolist(), cs[:nc.value].tolist()))
    return code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            # never built from here: a process that has touched the GPU must not
            # start other programs, so the build belongs to __graft_entry__.build()
            raise RuntimeError(f"{LIB} is missing: build it with make -C oracle")
        _lib = load(LIB)
    return _lib


def load(path):
    """A ctypes view of an oracle library built at path (the oracle, or a
    copy with a planted mistake that the oracle's own tests must catch)."""
    if True:
        L = C.CDLL(path)
        P = C.POINTER
        L.or_philox4x32_10.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        L.or_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int]
        L.or_uniform.restype = C.c_double
        L.or_gamma_stream.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double]
        L.or_gamma_stream.restype = C.c_double
        L.or_beta_stream.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double]
        L.or_beta_stream.restype = C.c_double
        L.or_parse_mod.argtypes = [C.c_char_p, P(Model), C.c_char_p, C.c_int]
        L.or_set_groups.argtypes = [P(Model), P(C.c_int), C.c_int, P(C.c_int), C.c_int]
        L.or_expm.argtypes = [P(C.c_double), C.c_double, P(C.c_double)]
        L.or_log_trans.argtypes = [P(Model), C.c_double, P(C.c_double)]
        L.or_encode.argtypes = [P(Model), P(Options), P(C.c_char_p), C.c_long, C.c_int]
        L.or_encode.restype = C.c_void_p
        L.or_free.argtypes = [C.c_void_p]
        L.or_elem_filtered.argtypes = [C.c_void_p]
        L.or_elem_info.argtypes = [C.c_void_p, P(C.c_int)]
        L.or_elem_arrays.argtypes = [C.c_void_p, P(C.c_int), P(C.c_int), P(C.c_double), P(C.c_int),
                                     P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_uint8)]
        L.or_loglik.argtypes = [P(Model), C.c_void_p, P(C.c_int), C.c_double, C.c_double]
        L.or_loglik.restype = C.c_double
        L.or_gbgc_rate.argtypes = [P(C.c_double), C.c_double, C.c_double, P(C.c_double)]
        L.or_gbgc_rate.restype = None
        L.or_loglik_gbgc.argtypes = [P(Model), C.c_void_p, P(C.c_int), P(C.c_int), P(C.c_double)]
        L.or_loglik_gbgc.restype = C.c_double
        L.or_run_element.argtypes = [P(Model), P(Options), C.c_void_p, P(ChainOut)]
        L.or_test_update_Z.argtypes = [C.c_int, P(C.c_int), C.c_int, P(C.c_int), P(C.c_int), P(C.c_double),
                                       C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int, C.c_int, P(C.c_int)]
        L.or_gt_loglik.argtypes = [P(Model), C.c_int, P(C.c_int), C.c_int, P(C.c_int), P(C.c_int), P(C.c_double),
                                   P(C.c_int), C.c_double, C.c_double, P(C.c_char_p), C.c_long, C.c_int, C.c_int, C.c_double]
        L.or_gt_loglik.restype = C.c_double
        L.or_gt_loglik_pi.argtypes = [P(Model), C.c_double] + L.or_gt_loglik.argtypes[1:]
        L.or_gt_loglik_pi.restype = C.c_double
        L.or_gt_sample_pi.argtypes = [P(Model), C.c_uint64, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                      P(C.c_double), P(C.c_double)] + L.or_gt_loglik.argtypes[1:]
        L.or_gt_sample_pi.restype = C.c_int
        L.or_wl_step.argtypes = [C.c_uint64, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double)]
        L.or_wl_step.restype = C.c_int
        L.or_test_eval2.argtypes = [P(Model), P(Options), C.c_void_p, C.c_int, P(C.c_double), P(C.c_double),
                                    P(C.c_double), P(C.c_byte), P(C.c_int)]
        L.or_test_eval2.restype = C.c_double
        L.or_test_sample_site.argtypes = [P(Model), C.c_void_p, P(C.c_int), C.c_double, C.c_double, C.c_uint64,
                                          C.c_int, C.c_int, C.c_int, P(C.c_int)]
        L.or_run_batch.argtypes = [P(Model), P(Options), P(C.c_void_p), C.c_int, C.c_int, P(C.c_double)]
        L.or_run_batch.restype = C.c_long
        PI, PD = P(C.c_int), P(C.c_double)
        L.or_gt_species.argtypes = [P(SpTree), PI, PD, C.c_int]
        L.or_gt_prior.argtypes = [P(SpTree), PI, PI, PD, C.c_int]
        L.or_gt_prior.restype = C.c_double
        L.or_gt_logpz.argtypes = [P(SpTree), PI, PD, PI, PI, C.c_double, C.c_double]
        L.or_gt_logpz.restype = C.c_double
        L.or_gt_pieces.argtypes = [P(SpTree), PI, PI, PD, C.c_int, PI, PI, PD]
        L.or_gt_site_logliks.argtypes = [P(Model), C.c_int, PI, C.c_int, C.c_int, PI, PI, PD, PI, C.c_double,
                                         C.c_double, P(C.c_char_p), C.c_long, C.c_int, C.c_int, C.c_double, PD]
        L.or_gt_site_logliks.restype = None
        L.or_gt_sample_history.argtypes = [P(Model), C.c_double, P(SpTree), PI, PI, PD, C.c_int, PI, C.c_double,
                                           C.c_double, C.c_uint64, C.c_int, C.c_int, P(C.c_char_p), C.c_long, C.c_int,
                                           C.c_int, C.c_double, P(C.c_int8)]
        L.or_gt_sample_history.restype = None
        L.or_gt_sample_tree.argtypes = [P(Model), C.c_double, P(SpTree), PI, PI, PD, PI, PI, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, PI, P(C.c_int8), PD, PD,
                                        C.c_uint64, C.c_int, C.c_int, P(C.c_char_p), C.c_long, C.c_int, C.c_int,
                                        C.c_double]
        L.or_test_sample_tree.argtypes = L.or_gt_sample_tree.argtypes + [C.c_int, PI, PD, PI]
        L.or_gt_branch_len.argtypes = [P(Model), C.c_double, P(SpTree), PI, PI, PD, PI, C.c_int, PI, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, PI, PD, PD,
                                       PD, C.c_uint64, C.c_int, C.c_int, P(C.c_char_p), C.c_long, C.c_int, C.c_int,
                                       C.c_double]
    return L


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def parse_mod(text: str):
    m = Model()
    names = C.create_string_buffer(64 * MAXN)
    rc = lib().or_parse_mod(text.encode(), C.byref(m), names, MAXN)
    if rc != 0:
        raise ValueError(f"or_parse_mod failed: {rc}")
    N = m.N
    nm = [names.raw[64 * i: 64 * (i + 1)].split(b"\0")[0].decode() for i in range(N)]
    return m, nm


def model_from_workload(w):
    m, names = parse_mod(w.mod_text)
    out = np.asarray(w.outgroup, np.int32)
    cons = np.asarray(w.conserve, np.int32)
    lib().or_set_groups(C.byref(m), _ptr(out, C.c_int), len(out), _ptr(cons, C.c_int), len(cons))
    return m


class Element:
    def __init__(self, model, opts, align, start, length):
        self.rows_keep = [np.ascontiguousarray(align[s]).tobytes() for s in range(align.shape[0])]
        rows = (C.c_char_p * len(self.rows_keep))(*self.rows_keep)
        self.model = model
        self.ptr = lib().or_encode(C.byref(model), C.byref(opts), rows, int(start), int(length))

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().or_free(self.ptr)
            self.ptr = None

    @property
    def filtered(self):
        return bool(lib().or_elem_filtered(self.ptr))

    def info(self):
        o = (C.c_int * 8)()
        lib().or_elem_info(self.ptr, o)
        return dict(GG=o[0], root=o[1], n_nodes=o[2], n_internal=o[3], S=o[4], N=o[5], filtered=o[6])

    def arrays(self):
        inf = self.info()
        N, GG, S = inf["N"], inf["GG"], inf["S"]
        ch = np.zeros(2 * N, np.int32)
        par = np.zeros(N, np.int32)
        d = np.zeros(N, np.float64)
        mis = np.zeros(N, np.int32)
        nodes = np.zeros(max(inf["n_nodes"], 1), np.int32)
        internal = np.zeros(max(inf["n_internal"], 1), np.int32)
        Tg = np.zeros(max(GG * N, 1), np.int32)
        codes = np.zeros(max(GG * S, 1), np.uint8)
        lib().or_elem_arrays(self.ptr, _ptr(ch, C.c_int), _ptr(par, C.c_int), _ptr(d, C.c_double),
                             _ptr(mis, C.c_int), _ptr(nodes, C.c_int), _ptr(internal, C.c_int),
                             _ptr(Tg, C.c_int), _ptr(codes, C.c_uint8))
        return dict(children=ch.reshape(N, 2), parent=par, dist=d, missing=mis,
                    nodes=nodes[: inf["n_nodes"]], internal=internal[: inf["n_internal"]],
                    Tg=Tg[: GG * N].reshape(GG, N), codes=codes[: GG * S].reshape(GG, S), **inf)

    def loglik_gbgc(self, Z, B, sel):
        Z = np.ascontiguousarray(Z, np.int32)
        B = np.ascontiguousarray(B, np.int32)
        sel = np.ascontiguousarray(sel, np.float64)
        return lib().or_loglik_gbgc(C.byref(self.model), self.ptr, _ptr(Z, C.c_int), _ptr(B, C.c_int),
                                    _ptr(sel, C.c_double))

    def loglik(self, Z, r1, r2):
        Z = np.ascontiguousarray(Z, np.int32)
        return lib().or_loglik(C.byref(self.model), self.ptr, _ptr(Z, C.c_int), float(r1), float(r2))

    def eval2(self, opts, ms, full, nrate, crate, Z, max_Z):
        """Eval2 of model slot ms over the given traces (T = burnin + mcmc)."""
        f = np.ascontiguousarray(full, np.float64)
        n = np.ascontiguousarray(nrate, np.float64)
        c = np.ascontiguousarray(crate, np.float64)
        z = np.ascontiguousarray(Z, np.int8)
        mz = np.ascontiguousarray(max_Z, np.int32)
        return lib().or_test_eval2(C.byref(self.model), C.byref(opts), self.ptr, int(ms), _ptr(f, C.c_double),
                                   _ptr(n, C.c_double), _ptr(c, C.c_double), _ptr(z, C.c_byte), _ptr(mz, C.c_int))

    def sample_site(self, Z, r1, r2, seed, m_sweep, ms=0, site=0):
        """The ancestral draw of one site at fixed Z and rates: every node's base."""
        z = np.ascontiguousarray(Z, np.int32)
        out = np.zeros(self.info()["N"], np.int32)
        rc = lib().or_test_sample_site(C.byref(self.model), self.ptr, _ptr(z, C.c_int), float(r1), float(r2),
                                       int(seed), int(m_sweep), int(ms), int(site), _ptr(out, C.c_int))
        if rc != 0:
            raise ValueError("or_test_sample_site failed")
        return out

    def run(self, opts, traces=True):
        """Run M0, M1, M2 (reference run order resZ 0, 2, 1)."""
        inf = self.info()
        T = opts.burnin + opts.mcmc
        N = inf["N"]
        outs = (ChainOut * 3)()
        keep = []
        for k in range(3):
            arrs = {}
            if traces:
                for f in ("loglik", "full", "nrate", "crate", "grate", "lrate", "l2rate"):
                    a = np.zeros(T, np.float64)
                    arrs[f] = a
                    setattr(outs[k], f, _ptr(a, C.c_double))
                z = np.zeros(T * N, np.int8)
                arrs["Z"] = z
                outs[k].Z = _ptr(z, C.c_byte)
            pz = np.zeros(N * 4, np.float64)
            arrs["postZ"] = pz
            outs[k].postZ = _ptr(pz, C.c_double)
            keep.append(arrs)
        rc = lib().or_run_element(C.byref(self.model), C.byref(opts), self.ptr, outs)
        if rc != 0:
            return None
        res = []
        for k in range(3):
            o = outs[k]
            d = dict(max_loglik=o.max_loglik, max_m=o.max_m, max_Z=np.array(o.max_Z[:N]),
                     vb=o.vb, med=np.array(o.med[:]), failure=o.failure,
                     accept_n=o.accept_n, accept_c=o.accept_c, flags=o.flags)
            for f, a in keep[k].items():
                d[f] = a.reshape(T, N) if f == "Z" else (a.reshape(N, 4) if f == "postZ" else a)
            res.append(d)
        return res


def run_batch(model, opts, elements, nthreads=0):
    arr = (C.c_void_p * len(elements))(*[e.ptr for e in elements])
    ml = np.zeros(3 * len(elements), np.float64)
    n = lib().or_run_batch(C.byref(model), C.byref(opts), arr, len(elements), nthreads, _ptr(ml, C.c_double))
    return n, ml.reshape(-1, 3)


def _gt_args(S, gchildren, seg_gene, seg_sp, seg_len, Z, r1, r2, align, start, length, gapchar, gap_prop):
    rows_keep = [np.ascontiguousarray(align[s]).tobytes() for s in range(align.shape[0])]
    rows = (C.c_char_p * len(rows_keep))(*rows_keep)
    keep = [rows_keep, np.ascontiguousarray(gchildren, np.int32).reshape(-1), np.ascontiguousarray(seg_gene, np.int32),
            np.ascontiguousarray(seg_sp, np.int32), np.ascontiguousarray(seg_len, np.float64),
            np.ascontiguousarray(Z, np.int32)]
    _, ch, sg, ss, sl, z = keep
    args = (int(S), _ptr(ch, C.c_int), len(sg), _ptr(sg, C.c_int), _ptr(ss, C.c_int), _ptr(sl, C.c_double),
            _ptr(z, C.c_int), float(r1), float(r2), rows, int(start), int(length), ord(gapchar), float(gap_prop))
    return args, keep


def gt_loglik_pi(model, a, S, gchildren, seg_gene, seg_sp, seg_len, Z, r1, r2, align, start, length, gapchar="-",
                 gap_prop=0.8):
    """or_gt_loglik with the element's base frequencies (a, 1/2 - a, 1/2 - a, a)."""
    args, keep = _gt_args(S, gchildren, seg_gene, seg_sp, seg_len, Z, r1, r2, align, start, length, gapchar, gap_prop)
    return lib().or_gt_loglik_pi(C.byref(model), float(a), *args)


def gt_sample_pi(model, seed, sweep, pi_delta, prior, resample, a, loglik, S, gchildren, seg_gene, seg_sp, seg_len,
                 Z, r1, r2, align, start, length, gapchar="-", gap_prop=0.8):
    """One sample_pi step; returns (a, loglik, accepted)."""
    args, keep = _gt_args(S, gchildren, seg_gene, seg_sp, seg_len, Z, r1, r2, align, start, length, gapchar, gap_prop)
    pa, ll = C.c_double(a), C.c_double(loglik)
    acc = lib().or_gt_sample_pi(C.byref(model), int(seed), int(sweep), float(pi_delta), float(prior[0]),
                                float(prior[1]), int(bool(resample)), C.byref(pa), C.byref(ll), *args)
    return pa.value, ll.value, acc


def wl_step(seed, m, loglik, logpz, state):
    """One Wang-Landau indicator step; state [7] updated in place."""
    ll = np.ascontiguousarray(loglik, np.float64)
    lz = np.ascontiguousarray(logpz, np.float64)
    assert state.dtype == np.float64 and state.flags.c_contiguous
    return lib().or_wl_step(int(seed), int(m), _ptr(ll, C.c_double), _ptr(lz, C.c_double), _ptr(state, C.c_double))


def gt_loglik(model, S, gchildren, seg_gene, seg_sp, seg_len, Z, r1, r2, align, start, length, gapchar="-",
              gap_prop=0.8):
    rows_keep = [np.ascontiguousarray(align[s]).tobytes() for s in range(align.shape[0])]
    rows = (C.c_char_p * len(rows_keep))(*rows_keep)
    ch = np.ascontiguousarray(gchildren, np.int32).reshape(-1)
    sg = np.ascontiguousarray(seg_gene, np.int32)
    ss = np.ascontiguousarray(seg_sp, np.int32)
    sl = np.ascontiguousarray(seg_len, np.float64)
    z = np.ascontiguousarray(Z, np.int32)
    return lib().or_gt_loglik(C.byref(model), int(S), _ptr(ch, C.c_int), len(sg), _ptr(sg, C.c_int),
                              _ptr(ss, C.c_int), _ptr(sl, C.c_double), _ptr(z, C.c_int), float(r1), float(r2),
                              rows, int(start), int(length), ord(gapchar), float(gap_prop))
