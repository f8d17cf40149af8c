"""Python binding of the engine's C ABI (include/phyloacc_gpu.h).

The engine mirrors PhyloAcc-ST's in-process interface: ``Engine`` plays the
role of ``BPP`` (tree, substitution model, priors, outgroup/conserve sets;
src/PhyloAcc-ST/bpp.hpp:173-404), ``add_elements`` the ``BPP_C`` constructor
(per-element encoding; src/PhyloAcc-ST/bpp_c.hpp:143-476), ``loglik`` the
fixed-(Z, r) likelihood ``BPP_C::log_f_Xz`` (src/PhyloAcc-ST/bpp_c2.cpp:392-425)
and ``run_chains`` the per-element loop body of main.cpp:452-500
(``initMCMC`` + ``Gibbs`` for M0, M1, M2).

All computation happens in libphyloacc_gpu.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libphyloacc_gpu.so")

PA_OK, PA_E_INPUT, PA_E_NUMERIC, PA_E_DEVICE, PA_E_OOM, PA_E_STATE = range(6)
PA_FP64, PA_FP32 = 0, 1
PA_RUN_BATCHED, PA_RUN_PER_ELEMENT = 0, 1
PA_ELEM_OK, PA_ELEM_FILTERED, PA_ELEM_NUMERIC_FAIL = 0, 1, 2
PA_FLAG_NAN, PA_FLAG_POS_INF, PA_FLAG_NEG_INF, PA_FLAG_SAMPLER, PA_FLAG_PROPOSAL_RANGE = 1, 2, 4, 8, 16

EXPORTS = ["pa_default_options", "pa_parse_mod", "pa_engine_create", "pa_engine_destroy", "pa_last_error",
           "pa_elements_add", "pa_element_get", "pa_loglik", "pa_run_chains", "pa_rng_uniform", "pa_device_count",
           "pa_chain_summary", "pa_elements_add_trees", "pa_gene_lengths", "pa_loglik_lengths",
           "pa_serve_create", "pa_serve_submit", "pa_serve_step", "pa_serve_poll", "pa_serve_get_stats",
           "pa_serve_destroy", "pa_serve_clock", "pa_set_tuning", "pa_get_tuning", "pa_autotune",
           "pa_serve_submit_lengths", "pa_release_scratch", "pa_loglik_gbgc",
           "pa_batch_create", "pa_batch_run", "pa_batch_destroy", "pa_loglik_last_stats",
           "pa_loglik_lengths_pi", "pa_sample_pi", "pa_loglik_lengths_split", "pa_wl_step",
           "pa_gt_prior", "pa_gt_logpz", "pa_gt_lengths", "pa_gt_branch_len", "pa_gt_sample_history",
           "pa_gt_sample_tree", "pa_gt_session_create", "pa_gt_session_branch_len", "pa_gt_session_sweep",
           "pa_gt_session_read", "pa_gt_session_destroy"]


class SpeciesTree(C.Structure):
    """pa_species_tree over numpy copies kept alive by the struct."""
    _fields_ = [("S", C.c_int32), ("N", C.c_int32), ("parent", C.POINTER(C.c_int32)),
                ("height", C.POINTER(C.c_double)), ("theta", C.POINTER(C.c_double))]

    @classmethod
    def make(cls, S, parent, height, theta):
        keep = [np.ascontiguousarray(parent, np.int32), np.ascontiguousarray(height, np.float64),
                np.ascontiguousarray(theta, np.float64)]
        t = cls(int(S), len(keep[0]), _p(keep[0], C.c_int32), _p(keep[1], C.c_double), _p(keep[2], C.c_double))
        t._keep = keep
        return t


class PaError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"phyloacc_gpu status {status}: {msg}")
        self.status = status


class Options(C.Structure):
    _fields_ = [
        ("burnin", C.c_int32), ("mcmc", C.c_int32), ("adapt_freq", C.c_int32),
        ("init_crate", C.c_double), ("init_nrate", C.c_double),
        ("cprior_a", C.c_double), ("cprior_b", C.c_double),
        ("nprior_a", C.c_double), ("nprior_b", C.c_double),
        ("ropt", C.c_int32), ("cub", C.c_double), ("nlb", C.c_double),
        ("init_grate", C.c_double), ("init_lrate", C.c_double), ("init_lrate2", C.c_double),
        ("hyper_grate_a", C.c_double), ("hyper_grate_b", C.c_double),
        ("hyper_lrate_a", C.c_double), ("hyper_lrate_b", C.c_double),
        ("hyper_lrate2_a", C.c_double), ("hyper_lrate2_b", C.c_double),
        ("gap_prop", C.c_double), ("conserve_prop", C.c_double), ("constomis", C.c_double),
        ("trim_gap_percent", C.c_double), ("min_len", C.c_int32), ("prune_tree", C.c_int32),
        ("gapchar", C.c_int32), ("seed", C.c_uint64), ("r1_shared_uniform", C.c_int32),
        ("compress_patterns", C.c_int32),
    ]


class Model(C.Structure):
    _fields_ = [
        ("S", C.c_int32), ("children", C.POINTER(C.c_int32)), ("parent", C.POINTER(C.c_int32)),
        ("dist", C.POINTER(C.c_double)), ("pi", C.POINTER(C.c_double)), ("Q", C.POINTER(C.c_double)),
        ("outgroup", C.POINTER(C.c_int32)), ("n_outgroup", C.c_int32),
        ("conserve", C.POINTER(C.c_int32)), ("n_conserve", C.c_int32),
    ]


class ElementInfo(C.Structure):
    _fields_ = [("filtered", C.c_int32), ("GG", C.c_int32), ("P", C.c_int32), ("root", C.c_int32),
                ("n_nodes", C.c_int32), ("n_internal", C.c_int32)]


class ChainResult(C.Structure):
    _fields_ = [("max_loglik", C.c_double), ("max_m", C.c_int32), ("status", C.c_int32),
                ("final_nrate", C.c_double), ("final_crate", C.c_double), ("final_grate", C.c_double),
                ("final_lrate", C.c_double), ("final_lrate2", C.c_double),
                ("accept_n", C.c_int32), ("accept_c", C.c_int32), ("flags", C.c_int32), ("pad", C.c_int32)]


class Traces(C.Structure):
    _fields_ = [(f, C.POINTER(C.c_double)) for f in ("loglik", "full", "nrate", "crate", "grate", "lrate", "l2rate")] + \
               [("Z", C.POINTER(C.c_int8)), ("max_Z", C.POINTER(C.c_int8))]


class RunStats(C.Structure):
    _fields_ = [("elapsed_ms", C.c_double), ("algorithmic_bytes", C.c_double), ("grid", C.c_int32),
                ("block_threads", C.c_int32), ("lds_bytes", C.c_int32), ("n_launches", C.c_int32),
                ("scratch_bytes", C.c_double)]


class RunConfig(C.Structure):
    _fields_ = [("precision", C.c_int32), ("mode", C.c_int32), ("block_threads", C.c_int32),
                ("keep_traces", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("block_threads", C.c_int32), ("shape", C.c_int32), ("round_nodes", C.c_int32), ("chunk_nodes", C.c_int32),
                ("rate", C.c_double)]


class ServeConfig(C.Structure):
    _fields_ = [("precision", C.c_int32), ("block_threads", C.c_int32), ("lanes", C.c_int32),
                ("max_chains", C.c_int32), ("hbm_budget_bytes", C.c_double)]


PA_SERVE_MAX_LANES, PA_SERVE_DEFAULT_LANES = 8, 3


PA_SERVE_CHAINS, PA_SERVE_LOGLIK = 0, 1


class ServeResult(C.Structure):
    _fields_ = [("ticket", C.c_int64), ("element", C.c_int32), ("status", C.c_int32), ("kind", C.c_int32),
                ("pad", C.c_int32), ("result", ChainResult * 3), ("loglik", C.c_double),
                ("arrival_s", C.c_double), ("admit_s", C.c_double), ("finish_s", C.c_double)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("admitted", C.c_int64), ("finished", C.c_int64),
                ("launches", C.c_int64), ("waiting", C.c_int32), ("in_flight", C.c_int32),
                ("scratch_bytes_in_use", C.c_double), ("device_ms", C.c_double),
                ("lanes_busy", C.c_int32), ("pad", C.c_int32)]


_lib = None


def lib():
    """Load the engine library; raises if the HIP build is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: build it with __graft_entry__.build() "
                               "(make -C phyloacc_amd/csrc); there is no CPU fallback")
        L = C.CDLL(LIBPATH)
        P = C.POINTER
        L.pa_default_options.argtypes = [P(Options)]
        L.pa_parse_mod.argtypes = [C.c_char_p, C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_int32),
                                   P(C.c_double), P(C.c_double), P(C.c_double), C.c_char_p]
        L.pa_engine_create.argtypes = [P(Model), P(Options), C.c_int32, P(C.c_void_p)]
        L.pa_engine_destroy.argtypes = [C.c_void_p]
        L.pa_last_error.argtypes = [C.c_void_p]
        L.pa_last_error.restype = C.c_char_p
        L.pa_elements_add.argtypes = [C.c_void_p, P(C.c_char_p), P(C.c_int64), P(C.c_int32), C.c_int32,
                                      P(C.c_int32), P(C.c_int32)]
        L.pa_element_get.argtypes = [C.c_void_p, C.c_int32, P(ElementInfo), P(C.c_int32), P(C.c_int32),
                                     P(C.c_double), P(C.c_int32), P(C.c_int32), P(C.c_uint8), P(C.c_int32),
                                     P(C.c_int32)]
        L.pa_loglik.argtypes = [C.c_void_p, P(C.c_int32), C.c_int32, P(C.c_int32), P(C.c_double), C.c_int32,
                                P(C.c_double)]
        L.pa_run_chains.argtypes = [C.c_void_p, P(C.c_int32), C.c_int32, P(RunConfig), P(ChainResult),
                                    P(Traces), P(RunStats)]
        L.pa_rng_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32]
        L.pa_rng_uniform.restype = C.c_double
        L.pa_chain_summary.argtypes = [C.c_void_p, C.c_int32, C.c_int32] + [P(C.c_double)] * 6 + \
            [P(C.c_int8), P(C.c_int8), C.c_int32, P(C.c_double), P(C.c_double), P(C.c_double)]
        L.pa_chain_summary.restype = C.c_int
        L.pa_elements_add_trees.argtypes = [C.c_void_p, P(C.c_char_p), P(C.c_int64), P(C.c_int32), C.c_int32,
                                            P(C.c_int32), P(C.c_int32), P(C.c_int32)]
        L.pa_elements_add_trees.restype = C.c_int
        L.pa_gene_lengths.argtypes = [C.c_int32, C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_double),
                                      P(C.c_int32), C.c_double, C.c_double, P(C.c_double)]
        L.pa_gene_lengths.restype = C.c_int
        L.pa_loglik_lengths.argtypes = [C.c_void_p, P(C.c_int32), C.c_int32, P(C.c_double), C.c_int32,
                                        P(C.c_double)]
        L.pa_loglik_lengths.restype = C.c_int
        L.pa_loglik_gbgc.argtypes = [C.c_void_p, P(C.c_int32), C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_double),
                                     C.c_int32, P(C.c_double)]
        L.pa_loglik_gbgc.restype = C.c_int
        L.pa_loglik_lengths_pi.argtypes = [C.c_void_p, P(C.c_int32), C.c_int32, P(C.c_double), P(C.c_double),
                                           C.c_int32, P(C.c_double)]
        L.pa_loglik_lengths_pi.restype = C.c_int
        L.pa_sample_pi.argtypes = [C.c_void_p, P(C.c_int32), C.c_int32, P(C.c_double), C.c_int32, C.c_double,
                                   C.c_double, C.c_double, C.c_int32, C.c_int32, P(C.c_double), P(C.c_double),
                                   P(C.c_int32)]
        L.pa_sample_pi.restype = C.c_int
        L.pa_loglik_lengths_split.argtypes = [C.c_void_p, P(C.c_int32), C.c_int32, P(C.c_double), P(C.c_double),
                                              P(C.c_int32), P(C.c_int32), C.c_int32, P(C.c_double), P(C.c_double)]
        L.pa_loglik_lengths_split.restype = C.c_int
        L.pa_wl_step.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(C.c_double), P(C.c_double), P(C.c_double),
                                 P(C.c_int32)]
        L.pa_wl_step.restype = C.c_int
        PI, PD = P(C.c_int32), P(C.c_double)
        L.pa_gt_prior.argtypes = [P(SpeciesTree), PI, PI, PD, PD]
        L.pa_gt_prior.restype = C.c_int
        L.pa_gt_logpz.argtypes = [P(SpeciesTree), PI, PI, PD, PI, PI, C.c_double, C.c_double, PD]
        L.pa_gt_logpz.restype = C.c_int
        L.pa_gt_lengths.argtypes = [P(SpeciesTree), PI, PI, PD, PI, C.c_double, C.c_double, PD]
        L.pa_gt_lengths.restype = C.c_int
        L.pa_gt_branch_len.argtypes = [C.c_void_p, P(SpeciesTree), PI, C.c_int32, PI, PI, PD, PI, PD, PD, PI, PI, PI,
                                       C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32, PD, PD,
                                       PD, PI]
        L.pa_gt_branch_len.restype = C.c_int
        L.pa_gt_sample_history.argtypes = [C.c_void_p, P(SpeciesTree), PI, C.c_int32, PI, PI, PD, PI, PD, PD, PI,
                                           C.c_int32, C.c_int32, P(C.c_int8), C.c_int32]
        L.pa_gt_sample_history.restype = C.c_int
        L.pa_gt_sample_tree.argtypes = [C.c_void_p, P(SpeciesTree), PI, C.c_int32, PI, PI, PD, PI, PD, PD, PI, PI,
                                        PI, P(C.c_int8), C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                        C.c_int32, C.c_int32, PD, PD, PI]
        L.pa_gt_sample_tree.restype = C.c_int
        L.pa_gt_session_create.argtypes = [C.c_void_p, P(SpeciesTree), PI, C.c_int32, PI, PI, PD, PI, PD, PD, PD, PD,
                                           PD, C.c_double, C.c_double, P(C.c_void_p)]
        L.pa_gt_session_create.restype = C.c_int
        L.pa_gt_session_branch_len.argtypes = [C.c_void_p, PI, PI, PI, C.c_double, C.c_int32, C.c_int32, C.c_int32, PI]
        L.pa_gt_session_branch_len.restype = C.c_int
        L.pa_gt_session_sweep.argtypes = [C.c_void_p, PI, C.c_int32, PI, PI, C.c_double, C.c_int32, C.c_int32,
                                          C.c_int32, PI, PD]
        L.pa_gt_session_sweep.restype = C.c_int
        L.pa_gt_session_read.argtypes = [C.c_void_p, PD, PD, PD, PD]
        L.pa_gt_session_read.restype = C.c_int
        L.pa_gt_session_destroy.argtypes = [C.c_void_p]
        L.pa_gt_session_destroy.restype = None
        L.pa_loglik_last_stats.argtypes = [C.c_void_p, P(RunStats)]
        L.pa_loglik_last_stats.restype = C.c_int
        L.pa_batch_create.argtypes = [C.c_void_p, P(C.c_int32), C.c_int32, P(RunConfig), P(C.c_void_p)]
        L.pa_batch_create.restype = C.c_int
        L.pa_batch_run.argtypes = [C.c_void_p, P(ChainResult), P(RunStats)]
        L.pa_batch_run.restype = C.c_int
        L.pa_batch_destroy.argtypes = [C.c_void_p]
        L.pa_batch_destroy.restype = None
        L.pa_set_tuning.argtypes = [C.c_void_p, P(Tuning)]
        L.pa_set_tuning.restype = C.c_int
        L.pa_get_tuning.argtypes = [C.c_void_p, P(Tuning)]
        L.pa_get_tuning.restype = C.c_int
        L.pa_autotune.argtypes = [C.c_void_p, P(C.c_int32), C.c_int32, C.c_int32, C.c_int32, C.c_int32, P(Tuning)]
        L.pa_autotune.restype = C.c_int
        L.pa_serve_create.argtypes = [C.c_void_p, P(ServeConfig), P(C.c_void_p)]
        L.pa_serve_create.restype = C.c_int
        L.pa_serve_submit.argtypes = [C.c_void_p, P(C.c_int32), C.c_int32, P(C.c_double), P(C.c_int64)]
        L.pa_serve_submit.restype = C.c_int
        L.pa_serve_submit_lengths.argtypes = [C.c_void_p, P(C.c_int32), C.c_int32, P(C.c_double), P(C.c_double),
                                              P(C.c_int64)]
        L.pa_serve_submit_lengths.restype = C.c_int
        L.pa_serve_step.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32)]
        L.pa_serve_step.restype = C.c_int
        L.pa_serve_poll.argtypes = [C.c_void_p, P(ServeResult), C.c_int32, P(C.c_int32)]
        L.pa_serve_poll.restype = C.c_int
        L.pa_serve_get_stats.argtypes = [C.c_void_p, P(ServeStats)]
        L.pa_serve_get_stats.restype = C.c_int
        L.pa_serve_destroy.argtypes = [C.c_void_p]
        L.pa_serve_destroy.restype = None
        L.pa_serve_clock.argtypes = []
        L.pa_serve_clock.restype = C.c_double
        L.pa_release_scratch.argtypes = [C.c_void_p]
        L.pa_release_scratch.restype = C.c_int
        L.pa_device_count.argtypes = []
        L.pa_device_count.restype = C.c_int32
        for f in ("pa_parse_mod", "pa_engine_create", "pa_elements_add", "pa_element_get", "pa_loglik",
                  "pa_run_chains"):
            getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def default_options(**kw) -> Options:
    o = Options()
    lib().pa_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def parse_mod(text: str, max_nodes: int = 512):
    """Reads a .mod file (LoadPhyloTree, src/PhyloAcc-ST/newick.cpp:431-517)."""
    S = C.c_int32()
    ch = np.zeros(2 * max_nodes, np.int32)
    par = np.zeros(max_nodes, np.int32)
    d = np.zeros(max_nodes, np.float64)
    pi = np.zeros(4, np.float64)
    Q = np.zeros(16, np.float64)
    names = C.create_string_buffer(64 * max_nodes)
    st = lib().pa_parse_mod(text.encode(), max_nodes, C.byref(S), _p(ch, C.c_int32), _p(par, C.c_int32),
                            _p(d, C.c_double), _p(pi, C.c_double), _p(Q, C.c_double), names)
    if st != PA_OK:
        raise PaError(st, "pa_parse_mod failed")
    N = 2 * S.value - 1
    nm = [names.raw[64 * i: 64 * (i + 1)].split(b"\0")[0].decode() for i in range(N)]
    return dict(S=S.value, N=N, children=ch[: 2 * N].reshape(N, 2).copy(), parent=par[:N].copy(),
                dist=d[:N].copy(), pi=pi, Q=Q.reshape(4, 4), names=nm)


class Engine:
    """One tree + substitution model + run options (BPP) and its encoded elements."""

    def __init__(self, S, children, parent, dist, pi, Q, outgroup=(), conserve=(), options: Options | None = None,
                 device: int = 0):
        self._keep = dict(
            children=np.ascontiguousarray(children, np.int32).reshape(-1),
            parent=np.ascontiguousarray(parent, np.int32),
            dist=np.ascontiguousarray(dist, np.float64),
            pi=np.ascontiguousarray(pi, np.float64),
            Q=np.ascontiguousarray(Q, np.float64).reshape(-1),
            out=np.ascontiguousarray(outgroup, np.int32),
            cons=np.ascontiguousarray(conserve, np.int32),
        )
        k = self._keep
        self.S = int(S)
        self.N = 2 * self.S - 1
        m = Model(self.S, _p(k["children"], C.c_int32), _p(k["parent"], C.c_int32), _p(k["dist"], C.c_double),
                  _p(k["pi"], C.c_double), _p(k["Q"], C.c_double), _p(k["out"], C.c_int32), len(k["out"]),
                  _p(k["cons"], C.c_int32), len(k["cons"]))
        self.options = options if options is not None else default_options()
        h = C.c_void_p()
        st = lib().pa_engine_create(C.byref(m), C.byref(self.options), int(device), C.byref(h))
        self._h = h.value
        if st != PA_OK:
            msg = lib().pa_last_error(self._h).decode() if self._h else "create failed"
            self.close()
            raise PaError(st, msg)
        self._rows = []

    @classmethod
    def from_workload(cls, w, options: Options | None = None, device: int = 0):
        t = parse_mod(w.mod_text)
        return cls(t["S"], t["children"], t["parent"], t["dist"], t["pi"], t["Q"], w.outgroup, w.conserve,
                   options, device)

    def close(self):
        if getattr(self, "_h", None):
            lib().pa_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, st):
        if st != PA_OK:
            raise PaError(st, lib().pa_last_error(self._h).decode())

    def add_elements(self, align: np.ndarray, starts, lengths):
        """Encode elements (BPP_C constructor).  align: [S, G] uint8 characters in tree leaf order."""
        align = np.ascontiguousarray(align, np.uint8)
        assert align.shape[0] == self.S
        self._rows.append(align)
        ptrs = (C.c_char_p * self.S)(*[C.c_char_p(align[s].ctypes.data) for s in range(self.S)])
        st = np.ascontiguousarray(starts, np.int64)
        ln = np.ascontiguousarray(lengths, np.int32)
        n = len(st)
        ids = np.zeros(n, np.int32)
        status = np.zeros(n, np.int32)
        self._check(lib().pa_elements_add(self._h, ptrs, _p(st, C.c_int64), _p(ln, C.c_int32), n,
                                          _p(ids, C.c_int32), _p(status, C.c_int32)))
        return ids, status

    def add_elements_with_trees(self, align: np.ndarray, starts, lengths, children):
        """Elements carrying their own tree (gene trees): children [n, N, 2]."""
        align = np.ascontiguousarray(align, np.uint8)
        self._rows.append(align)
        ptrs = (C.c_char_p * self.S)(*[C.c_char_p(align[s].ctypes.data) for s in range(self.S)])
        st = np.ascontiguousarray(starts, np.int64)
        ln = np.ascontiguousarray(lengths, np.int32)
        ch = np.ascontiguousarray(children, np.int32)
        n = len(st)
        ids = np.zeros(n, np.int32)
        status = np.zeros(n, np.int32)
        self._check(lib().pa_elements_add_trees(self._h, ptrs, _p(st, C.c_int64), _p(ln, C.c_int32), n,
                                                _p(ch, C.c_int32), _p(ids, C.c_int32), _p(status, C.c_int32)))
        return ids, status

    def loglik_lengths(self, ids, lengths, precision=PA_FP64):
        ids = np.ascontiguousarray(ids, np.int32)
        L = np.ascontiguousarray(lengths, np.float64).reshape(len(ids), self.N)
        out = np.zeros(len(ids), np.float64)
        self._check(lib().pa_loglik_lengths(self._h, _p(ids, C.c_int32), len(ids), _p(L, C.c_double),
                                            int(precision), _p(out, C.c_double)))
        return out

    def loglik_gbgc(self, ids, Z, B, sel, precision=PA_FP64):
        """gBGC model: Z, B [n, N]; sel [n, 3] = (conserved selection,
        accelerated selection, gBGC bias)."""
        ids = np.ascontiguousarray(ids, np.int32)
        Z = np.ascontiguousarray(Z, np.int32).reshape(len(ids), self.N)
        B = np.ascontiguousarray(B, np.int32).reshape(len(ids), self.N)
        sel = np.ascontiguousarray(sel, np.float64).reshape(len(ids), 3)
        out = np.zeros(len(ids), np.float64)
        self._check(lib().pa_loglik_gbgc(self._h, _p(ids, C.c_int32), len(ids), _p(Z, C.c_int32), _p(B, C.c_int32),
                                         _p(sel, C.c_double), int(precision), _p(out, C.c_double)))
        return out

    def loglik_lengths_pi(self, ids, lengths, piA, precision=PA_FP64):
        """Gene-tree log-likelihoods with per-element strand-symmetric base
        frequencies (a, 1/2 - a, 1/2 - a, a), piA[n] = a."""
        ids = np.ascontiguousarray(ids, np.int32)
        L = np.ascontiguousarray(lengths, np.float64).reshape(len(ids), self.N)
        a = np.ascontiguousarray(piA, np.float64)
        out = np.zeros(len(ids), np.float64)
        self._check(lib().pa_loglik_lengths_pi(self._h, _p(ids, C.c_int32), len(ids), _p(L, C.c_double),
                                               _p(a, C.c_double), int(precision), _p(out, C.c_double)))
        return out

    def sample_pi(self, ids, lengths, m, piA, loglik, pi_delta=0.1, prior=(1.0, 1.0), resample=False,
                  precision=PA_FP64):
        """One batched sample_pi step; returns (piA, loglik, accepted) after it."""
        ids = np.ascontiguousarray(ids, np.int32)
        L = np.ascontiguousarray(lengths, np.float64).reshape(len(ids), self.N)
        a = np.array(piA, np.float64)
        ll = np.array(loglik, np.float64)
        acc = np.zeros(len(ids), np.int32)
        self._check(lib().pa_sample_pi(self._h, _p(ids, C.c_int32), len(ids), _p(L, C.c_double), int(m),
                                       float(pi_delta), float(prior[0]), float(prior[1]), int(bool(resample)),
                                       int(precision), _p(a, C.c_double), _p(ll, C.c_double), _p(acc, C.c_int32)))
        return a, ll, acc

    def loglik_lengths_split(self, ids, lengths, lens0, lens1, piA=None, precision=PA_FP64):
        """Gene-tree log-likelihoods over each element's first lens0 and first
        lens1 sites; returns (out0, out1)."""
        ids = np.ascontiguousarray(ids, np.int32)
        L = np.ascontiguousarray(lengths, np.float64).reshape(len(ids), self.N)
        l0 = np.ascontiguousarray(lens0, np.int32)
        l1 = np.ascontiguousarray(lens1, np.int32)
        a = None if piA is None else np.ascontiguousarray(piA, np.float64)
        o0 = np.zeros(len(ids), np.float64)
        o1 = np.zeros(len(ids), np.float64)
        self._check(lib().pa_loglik_lengths_split(self._h, _p(ids, C.c_int32), len(ids), _p(L, C.c_double),
                                                  None if a is None else _p(a, C.c_double), _p(l0, C.c_int32),
                                                  _p(l1, C.c_int32), int(precision), _p(o0, C.c_double),
                                                  _p(o1, C.c_double)))
        return o0, o1

    def wl_step(self, m, loglik, logpz, state):
        """One Wang-Landau indicator step; loglik, logpz [n, 2], state [n, 7]
        (updated in place).  Returns the indicators."""
        ll = np.ascontiguousarray(loglik, np.float64).reshape(-1, 2)
        lz = np.ascontiguousarray(logpz, np.float64).reshape(-1, 2)
        assert state.dtype == np.float64 and state.flags.c_contiguous and state.shape == (len(ll), 7)
        ind = np.zeros(len(ll), np.int32)
        self._check(lib().pa_wl_step(self._h, len(ll), int(m), _p(ll, C.c_double), _p(lz, C.c_double),
                                     _p(state, C.c_double), _p(ind, C.c_int32)))
        return ind

    def gt_branch_len(self, sp, ids, children, parent, heights, Z, rates, gnode, indicator, lens, loglik, logpz,
                      prior, m, slot, delta=1.0, cons_to_mis=0.01, ncons_to_mis=0.5, piA=None, precision=PA_FP64):
        """One batched Sample_BranchLen move.  heights [n, N], loglik, logpz
        [n, 2] and prior [n] (float64, C order) are updated in place; returns
        accepted [n]."""
        ids = np.ascontiguousarray(ids, np.int32)
        n = len(ids)
        ch = np.ascontiguousarray(children, np.int32).reshape(n, self.N, 2)
        par = np.ascontiguousarray(parent, np.int32).reshape(n, self.N)
        for a in (heights, loglik, logpz, prior):
            assert a.dtype == np.float64 and a.flags.c_contiguous
        assert heights.shape == (n, self.N) and loglik.shape == (n, 2) and logpz.shape == (n, 2) and prior.shape == (n,)
        z = np.ascontiguousarray(Z, np.int32).reshape(n, self.N)
        r = np.ascontiguousarray(rates, np.float64).reshape(n, 2)
        g = np.ascontiguousarray(np.broadcast_to(gnode, (n,)), np.int32)
        ind = np.ascontiguousarray(np.broadcast_to(indicator, (n,)), np.int32)
        ln = np.ascontiguousarray(np.broadcast_to(lens, (n, 2)), np.int32)
        a = None if piA is None else np.ascontiguousarray(piA, np.float64)
        acc = np.zeros(n, np.int32)
        self._check(lib().pa_gt_branch_len(
            self._h, C.byref(sp), _p(ids, C.c_int32), n, _p(ch, C.c_int32), _p(par, C.c_int32),
            _p(heights, C.c_double), _p(z, C.c_int32), _p(r, C.c_double), None if a is None else _p(a, C.c_double),
            _p(g, C.c_int32), _p(ind, C.c_int32), _p(ln, C.c_int32), float(delta), float(cons_to_mis),
            float(ncons_to_mis), int(m), int(slot), int(precision), _p(loglik, C.c_double), _p(logpz, C.c_double),
            _p(prior, C.c_double), _p(acc, C.c_int32)))
        return acc

    def gt_sample_history(self, sp, ids, children, parent, heights, Z, rates, length, m, piA=None,
                          precision=PA_FP64, stride=None, to_host=True):
        """Batched Update_Tg(tosample=true): the histories of the first
        length[i] sites stay on the device for the next gt_sample_tree of the
        same ids; with to_host also returned as int8 [n, N, stride]."""
        ids = np.ascontiguousarray(ids, np.int32)
        n = len(ids)
        ch = np.ascontiguousarray(children, np.int32).reshape(n, self.N, 2)
        par = np.ascontiguousarray(parent, np.int32).reshape(n, self.N)
        h = np.ascontiguousarray(heights, np.float64).reshape(n, self.N)
        z = np.ascontiguousarray(Z, np.int32).reshape(n, self.N)
        r = np.ascontiguousarray(rates, np.float64).reshape(n, 2)
        ln = np.ascontiguousarray(np.broadcast_to(length, (n,)), np.int32)
        stride = int(max(1, ln.max() if n else 1) if stride is None else stride)
        a = None if piA is None else np.ascontiguousarray(piA, np.float64)
        out = np.zeros((n, self.N, stride), np.int8) if to_host else None
        self._check(lib().pa_gt_sample_history(
            self._h, C.byref(sp), _p(ids, C.c_int32), n, _p(ch, C.c_int32), _p(par, C.c_int32), _p(h, C.c_double),
            _p(z, C.c_int32), _p(r, C.c_double), None if a is None else _p(a, C.c_double), _p(ln, C.c_int32), int(m),
            int(precision), None if out is None else _p(out, C.c_int8), stride))
        return out

    def gt_sample_tree(self, sp, ids, children, parent, heights, Z, rates, branch, indicator, lens, Tg, loglik,
                       logpz, m, slot, cons_to_mis=0.01, ncons_to_mis=0.5, root_ingrp=None, piA=None,
                       precision=PA_FP64):
        """One batched Sample_tree2 move.  children [n, N, 2], parent,
        heights [n, N], loglik, logpz [n, 2] are updated in place (C order;
        int32 / float64); Tg [n, N, stride] as gt_sample_history returns it,
        or None for the device histories of the last gt_sample_history of
        these ids.  Returns accepted [n]."""
        ids = np.ascontiguousarray(ids, np.int32)
        n = len(ids)
        for a, t in ((children, np.int32), (parent, np.int32), (heights, np.float64), (loglik, np.float64),
                     (logpz, np.float64)):
            assert a.dtype == t and a.flags.c_contiguous
        assert children.shape == (n, self.N, 2) and parent.shape == (n, self.N) and heights.shape == (n, self.N)
        z = np.ascontiguousarray(Z, np.int32).reshape(n, self.N)
        r = np.ascontiguousarray(rates, np.float64).reshape(n, 2)
        b = np.ascontiguousarray(np.broadcast_to(branch, (n,)), np.int32)
        ind = np.ascontiguousarray(np.broadcast_to(indicator, (n,)), np.int32)
        ln = np.ascontiguousarray(np.broadcast_to(lens, (n, 2)), np.int32)
        T = None if Tg is None else np.ascontiguousarray(Tg, np.int8)
        assert T is None or T.shape[:2] == (n, self.N)
        a = None if piA is None else np.ascontiguousarray(piA, np.float64)
        acc = np.zeros(n, np.int32)
        ri = self.N - 1 if root_ingrp is None else int(root_ingrp)
        self._check(lib().pa_gt_sample_tree(
            self._h, C.byref(sp), _p(ids, C.c_int32), n, _p(children, C.c_int32), _p(parent, C.c_int32),
            _p(heights, C.c_double), _p(z, C.c_int32), _p(r, C.c_double), None if a is None else _p(a, C.c_double),
            _p(b, C.c_int32), _p(ind, C.c_int32), _p(ln, C.c_int32), None if T is None else _p(T, C.c_int8),
            0 if T is None else T.shape[2],
            float(cons_to_mis), float(ncons_to_mis), ri, int(m), int(slot), int(precision), _p(loglik, C.c_double),
            _p(logpz, C.c_double), _p(acc, C.c_int32)))
        return acc

    def loglik_stats(self):
        """Kernel time and counted bytes of the last likelihood call."""
        st = RunStats()
        self._check(lib().pa_loglik_last_stats(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in RunStats._fields_}

    def release_scratch(self):
        """Free the device scratch the likelihood calls keep between calls."""
        self._check(lib().pa_release_scratch(self._h))

    def element(self, eid: int):
        info = ElementInfo()
        self._check(lib().pa_element_get(self._h, int(eid), C.byref(info), None, None, None, None, None, None,
                                         None, None))
        N, S = self.N, self.S
        d = dict(filtered=info.filtered, GG=info.GG, P=info.P, root=info.root, n_nodes=info.n_nodes,
                 n_internal=info.n_internal)
        if info.filtered:
            return d
        ch = np.zeros(2 * N, np.int32)
        par = np.zeros(N, np.int32)
        dist = np.zeros(N, np.float64)
        mis = np.zeros(N, np.int32)
        nodes = np.zeros(N, np.int32)
        codes = np.zeros(info.GG * S, np.uint8)
        site = np.zeros(info.GG, np.int32)
        w = np.zeros(max(info.P, 1), np.int32)
        self._check(lib().pa_element_get(self._h, int(eid), C.byref(info), _p(ch, C.c_int32), _p(par, C.c_int32),
                                         _p(dist, C.c_double), _p(mis, C.c_int32), _p(nodes, C.c_int32),
                                         _p(codes, C.c_uint8), _p(site, C.c_int32), _p(w, C.c_int32)))
        d.update(children=ch.reshape(N, 2), parent=par, dist=dist, missing=mis, nodes=nodes[: info.n_nodes],
                 codes=codes.reshape(info.GG, S), site_pattern=site, weights=w[: info.P])
        return d

    def autotune(self, ids, precision=PA_FP32, sample=0, states=8):
        """Time every block size (64/128/256) and pattern-loop shape on the
        batch (sample > 0: on that many elements evenly spaced by pattern
        count) with short chains; the fastest pair is kept for
        run_chains(block_threads=0) and Batch(block_threads=0)."""
        ids = np.ascontiguousarray(ids, np.int32)
        t = Tuning()
        self._check(lib().pa_autotune(self._h, _p(ids, C.c_int32), len(ids), int(precision), int(sample),
                                      int(states), C.byref(t)))
        return {f: getattr(t, f) for f, _ in Tuning._fields_}

    def set_tuning(self, block_threads=0, shape=0):
        self._check(lib().pa_set_tuning(self._h, C.byref(Tuning(block_threads, shape, 0, 0, 0.0))))

    def tuning(self):
        t = Tuning()
        self._check(lib().pa_get_tuning(self._h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in Tuning._fields_}

    def loglik(self, ids, Z, rates, precision=PA_FP64):
        ids = np.ascontiguousarray(ids, np.int32)
        Z = np.ascontiguousarray(Z, np.int32).reshape(len(ids), self.N)
        rates = np.ascontiguousarray(rates, np.float64).reshape(len(ids), 2)
        out = np.zeros(len(ids), np.float64)
        self._check(lib().pa_loglik(self._h, _p(ids, C.c_int32), len(ids), _p(Z, C.c_int32),
                                    _p(rates, C.c_double), int(precision), _p(out, C.c_double)))
        return out

    def chain_summary(self, eid, model, traces, precision=PA_FP64, want_vb=True):
        """Eval2 marginal likelihood, Output_init medians and Z frequencies of one
        element's model chain from its traces (dict of [T] arrays and Z [T, N])."""
        T = self.options.burnin + self.options.mcmc
        arr = {f: np.ascontiguousarray(traces[f], np.float64).reshape(T)
               for f in ("full", "nrate", "crate", "grate", "lrate", "l2rate")}
        Z = np.ascontiguousarray(traces["Z"], np.int8).reshape(T, self.N)
        mz = np.ascontiguousarray(traces["max_Z"], np.int8).reshape(self.N)
        vb = np.zeros(1)
        med = np.zeros(5)
        pz = np.zeros(self.N * 4)
        self._check(lib().pa_chain_summary(
            self._h, int(eid), int(model), *[_p(arr[f], C.c_double) for f in ("full", "nrate", "crate", "grate",
                                                                                 "lrate", "l2rate")],
            _p(Z, C.c_int8), _p(mz, C.c_int8), int(precision), _p(vb, C.c_double) if want_vb else None,
            _p(med, C.c_double), _p(pz, C.c_double)))
        return dict(vb=float(vb[0]) if want_vb else None, med=med, postZ=pz.reshape(self.N, 4))

    def run_chains(self, ids, precision=PA_FP32, mode=PA_RUN_BATCHED, traces=False, block_threads=0,
                   keep_traces=False):
        """Runs M0, M1, M2 for every id.  Returns per-(element, model) arrays, models in run order."""
        ids = np.ascontiguousarray(ids, np.int32)
        n = len(ids)
        T = self.options.burnin + self.options.mcmc
        N = self.N
        res = (ChainResult * (3 * n))()
        tr = Traces()
        out = {}
        maxZ = np.zeros((n, 3, N), np.int8)
        tr.max_Z = _p(maxZ, C.c_int8)
        if traces:
            for f in ("loglik", "full", "nrate", "crate", "grate", "lrate", "l2rate"):
                a = np.zeros((n, 3, T), np.float64)
                out[f] = a
                setattr(tr, f, _p(a, C.c_double))
            z = np.zeros((n, 3, T, N), np.int8)
            out["Z"] = z
            tr.Z = _p(z, C.c_int8)
        cfg = RunConfig(int(precision), int(mode), int(block_threads), int(bool(keep_traces)))
        stats = RunStats()
        self._check(lib().pa_run_chains(self._h, _p(ids, C.c_int32), n, C.byref(cfg), res, C.byref(tr),
                                        C.byref(stats)))
        out["max_Z"] = maxZ
        out["elapsed_ms"] = stats.elapsed_ms
        out["stats"] = {f: getattr(stats, f) for f, _ in RunStats._fields_}
        for f, t in (("max_loglik", np.float64), ("max_m", np.int32), ("status", np.int32),
                     ("final_nrate", np.float64), ("final_crate", np.float64), ("final_grate", np.float64),
                     ("final_lrate", np.float64), ("final_lrate2", np.float64), ("accept_n", np.int32),
                     ("accept_c", np.int32), ("flags", np.int32)):
            out[f] = np.array([getattr(res[i], f) for i in range(3 * n)], t).reshape(n, 3)
        return out


RESULT_FIELDS = (("max_loglik", np.float64), ("max_m", np.int32), ("status", np.int32),
                 ("final_nrate", np.float64), ("final_crate", np.float64), ("final_grate", np.float64),
                 ("final_lrate", np.float64), ("final_lrate2", np.float64), ("accept_n", np.int32),
                 ("accept_c", np.int32), ("flags", np.int32))


class Batch:
    """A batched chain run prepared once and captured as a HIP graph;
    ``run()`` replays it (same results as ``Engine.run_chains`` on the ids)."""

    def __init__(self, engine: "Engine", ids, precision=PA_FP32, block_threads=0, keep_traces=False):
        self.engine = engine
        self.ids = np.ascontiguousarray(ids, np.int32)
        cfg = RunConfig(int(precision), int(PA_RUN_BATCHED), int(block_threads), int(bool(keep_traces)))
        h = C.c_void_p()
        engine._check(lib().pa_batch_create(engine._h, _p(self.ids, C.c_int32), len(self.ids), C.byref(cfg),
                                            C.byref(h)))
        self._h = h

    def run(self):
        n = len(self.ids)
        res = (ChainResult * (3 * n))()
        stats = RunStats()
        self.engine._check(lib().pa_batch_run(self._h, res, C.byref(stats)))
        out = {"elapsed_ms": stats.elapsed_ms, "stats": {f: getattr(stats, f) for f, _ in RunStats._fields_}}
        for f, t in RESULT_FIELDS:
            out[f] = np.array([getattr(res[i], f) for i in range(3 * n)], t).reshape(n, 3)
        return out

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self.engine, "_h", None):    # a closed engine has freed its batches
                lib().pa_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def gene_lengths(N, seg_gene, seg_species, seg_len, Z, r1, r2):
    sg = np.ascontiguousarray(seg_gene, np.int32)
    ss = np.ascontiguousarray(seg_species, np.int32)
    sl = np.ascontiguousarray(seg_len, np.float64)
    z = np.ascontiguousarray(Z, np.int32)
    out = np.zeros(N, np.float64)
    st = lib().pa_gene_lengths(int(N), len(sg), _p(sg, C.c_int32), _p(ss, C.c_int32), _p(sl, C.c_double),
                               _p(z, C.c_int32), float(r1), float(r2), _p(out, C.c_double))
    if st != PA_OK:
        raise PaError(st, "pa_gene_lengths failed")
    return out


class GtSession:
    """Device-resident gene trees for batched node-height moves
    (pa_gt_session_*): created from the trees and chain state, moved on the
    device, read back at the end."""

    def __init__(self, eng, sp, ids, children, parent, heights, Z, rates, loglik, logpz, prior, piA=None,
                 cons_to_mis=0.01, ncons_to_mis=0.5):
        self.eng = eng
        ids = np.ascontiguousarray(ids, np.int32)
        self.n, self.N = len(ids), eng.N
        arrs = [np.ascontiguousarray(children, np.int32), np.ascontiguousarray(parent, np.int32),
                np.ascontiguousarray(heights, np.float64), np.ascontiguousarray(Z, np.int32),
                np.ascontiguousarray(rates, np.float64), np.ascontiguousarray(loglik, np.float64),
                np.ascontiguousarray(logpz, np.float64), np.ascontiguousarray(prior, np.float64)]
        a = None if piA is None else np.ascontiguousarray(piA, np.float64)
        h = C.c_void_p()
        eng._check(lib().pa_gt_session_create(
            eng._h, C.byref(sp), _p(ids, C.c_int32), self.n, _p(arrs[0], C.c_int32), _p(arrs[1], C.c_int32),
            _p(arrs[2], C.c_double), _p(arrs[3], C.c_int32), _p(arrs[4], C.c_double),
            None if a is None else _p(a, C.c_double), _p(arrs[5], C.c_double), _p(arrs[6], C.c_double),
            _p(arrs[7], C.c_double), float(cons_to_mis), float(ncons_to_mis), C.byref(h)))
        self._h = h
        self._sp = sp

    def branch_len(self, gnode, indicator, lens, m, slot, delta=1.0, precision=PA_FP64, want_accepted=True):
        g = np.ascontiguousarray(np.broadcast_to(gnode, (self.n,)), np.int32)
        ind = np.ascontiguousarray(np.broadcast_to(indicator, (self.n,)), np.int32)
        ln = np.ascontiguousarray(np.broadcast_to(lens, (self.n, 2)), np.int32)
        acc = np.zeros(self.n, np.int32) if want_accepted else None
        self.eng._check(lib().pa_gt_session_branch_len(self._h, _p(g, C.c_int32), _p(ind, C.c_int32),
                                                       _p(ln, C.c_int32), float(delta), int(m), int(slot),
                                                       int(precision), None if acc is None else _p(acc, C.c_int32)))
        return acc

    def sweep(self, gnodes, indicator, lens, m, slot0, delta=1.0, precision=PA_FP64):
        """k moves in a row (gnodes [k, n], or [k] for the same node in
        every element), queued at once; returns (accepted [k, n], the
        likelihood launches' summed ms)."""
        g = np.asarray(gnodes, np.int32)
        if g.ndim == 1:
            g = np.repeat(g[:, None], self.n, axis=1)
        g = np.ascontiguousarray(g, np.int32)
        k = g.shape[0]
        ind = np.ascontiguousarray(np.broadcast_to(indicator, (self.n,)), np.int32)
        ln = np.ascontiguousarray(np.broadcast_to(lens, (self.n, 2)), np.int32)
        acc = np.zeros((k, self.n), np.int32)
        ms = C.c_double(0)
        self.eng._check(lib().pa_gt_session_sweep(self._h, _p(g, C.c_int32), k, _p(ind, C.c_int32),
                                                  _p(ln, C.c_int32), float(delta), int(m), int(slot0),
                                                  int(precision), _p(acc, C.c_int32), C.byref(ms)))
        return acc, ms.value

    def read(self):
        """(heights [n, N], loglik [n, 2], logpz [n, 2], prior [n])."""
        h = np.zeros((self.n, self.N))
        ll = np.zeros((self.n, 2))
        lz = np.zeros((self.n, 2))
        pr = np.zeros(self.n)
        self.eng._check(lib().pa_gt_session_read(self._h, _p(h, C.c_double), _p(ll, C.c_double),
                                                 _p(lz, C.c_double), _p(pr, C.c_double)))
        return h, ll, lz, pr

    def close(self):
        if self._h:
            lib().pa_gt_session_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _gt_call(name, sp, children, parent, heights, *rest):
    ch = np.ascontiguousarray(children, np.int32).reshape(-1)
    par = np.ascontiguousarray(parent, np.int32)
    h = np.ascontiguousarray(heights, np.float64)
    st = getattr(lib(), name)(C.byref(sp), _p(ch, C.c_int32), _p(par, C.c_int32), _p(h, C.c_double), *rest)
    if st != PA_OK:
        raise PaError(st, f"{name} failed")


def gt_prior(sp, children, parent, heights):
    """GTree::priorTree of one gene tree (host)."""
    out = C.c_double()
    _gt_call("pa_gt_prior", sp, children, parent, heights, C.byref(out))
    return out.value


def gt_logpz(sp, children, parent, heights, leaf_missing, Z, cons_to_mis, ncons_to_mis):
    """GTree::get_logpZ1 of one gene tree (host)."""
    lm = np.ascontiguousarray(leaf_missing, np.int32)
    z = np.ascontiguousarray(Z, np.int32)
    out = C.c_double()
    _gt_call("pa_gt_logpz", sp, children, parent, heights, _p(lm, C.c_int32), _p(z, C.c_int32),
             float(cons_to_mis), float(ncons_to_mis), C.byref(out))
    return out.value


def gt_lengths(sp, children, parent, heights, Z, r1, r2):
    """Substitution length of every gene branch of one gene tree (host)."""
    z = np.ascontiguousarray(Z, np.int32)
    out = np.zeros(sp.N, np.float64)
    _gt_call("pa_gt_lengths", sp, children, parent, heights, _p(z, C.c_int32), float(r1), float(r2),
             _p(out, C.c_double))
    return out


def device_count() -> int:
    return int(lib().pa_device_count())


def rng_uniform(seed, c0, c1, c2, c3, lane):
    return lib().pa_rng_uniform(seed, c0, c1, c2, c3, lane)


class Server:
    """Continuous request loop over an engine's encoded elements: up to
    ``lanes`` launches in flight, each holding part of the HBM pool; every
    element retires as soon as its own chains end."""

    def __init__(self, engine: "Engine", precision=PA_FP32, block_threads=0, lanes=0, max_chains=0,
                 hbm_budget_bytes=0.0):
        self.engine = engine
        cfg = ServeConfig(int(precision), int(block_threads), int(lanes), int(max_chains),
                          float(hbm_budget_bytes))
        h = C.c_void_p()
        engine._check(lib().pa_serve_create(engine._h, C.byref(cfg), C.byref(h)))
        self._h = h

    def _check(self, st):
        if st != PA_OK:
            raise PaError(st, lib().pa_last_error(self.engine._h).decode())

    def submit(self, ids, arrival_s=None):
        ids = np.ascontiguousarray(ids, np.int32)
        t = np.zeros(len(ids), np.int64)
        arr = None if arrival_s is None else np.ascontiguousarray(arrival_s, np.float64)
        self._check(lib().pa_serve_submit(self._h, _p(ids, C.c_int32), len(ids),
                                          None if arr is None else _p(arr, C.c_double), _p(t, C.c_int64)))
        return t

    def submit_lengths(self, ids, lengths, arrival_s=None):
        """Log-likelihood requests for own-tree elements at branch lengths [n, N]."""
        ids = np.ascontiguousarray(ids, np.int32)
        L = np.ascontiguousarray(lengths, np.float64).reshape(len(ids), -1)
        t = np.zeros(len(ids), np.int64)
        arr = None if arrival_s is None else np.ascontiguousarray(arrival_s, np.float64)
        self._check(lib().pa_serve_submit_lengths(self._h, _p(ids, C.c_int32), len(ids), _p(L, C.c_double),
                                                  None if arr is None else _p(arr, C.c_double), _p(t, C.c_int64)))
        return t

    def step(self, wait=True) -> int:
        n = C.c_int32()
        self._check(lib().pa_serve_step(self._h, int(bool(wait)), C.byref(n)))
        return n.value

    def poll(self, max_results=4096):
        buf = (ServeResult * max_results)()
        n = C.c_int32()
        self._check(lib().pa_serve_poll(self._h, buf, max_results, C.byref(n)))
        return [buf[i] for i in range(n.value)]

    def stats(self) -> ServeStats:
        s = ServeStats()
        self._check(lib().pa_serve_get_stats(self._h, C.byref(s)))
        return s

    def close(self):
        if getattr(self, "_h", None):
            lib().pa_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def serve_clock() -> float:
    return lib().pa_serve_clock()
