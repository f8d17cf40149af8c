"""The allocation-free likelihood entry point (pa_loglik_noalloc) against
pa_loglik, through the C ABI: device-resident ids, Z, rates and outputs,
two calls in a row on one engine (the second smaller), each compared with
pa_loglik on the same input at both precisions.

Device buffers come from the HIP runtime the engine library already loaded
(plain hipMalloc / hipMemcpy through ctypes); the entry point itself
allocates nothing.
"""
import ctypes as C

import numpy as np
import pytest

from phyloacc_amd import engine as E
from phyloacc_amd import synth

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    E.lib()                                   # the engine library brings in the HIP runtime
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class _Dev:
    """One device buffer holding a copy of a host array."""

    def __init__(self, hip, a):
        self.hip, self.n = hip, a.nbytes
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(self.n, 8)) == 0
        if self.n:
            assert hip.hipMemcpy(self.p, a.ctypes.data, self.n, H2D) == 0

    def get(self, dtype, count):
        out = np.zeros(count, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(eng, hip, ids, Z, rates, prec):
    L = E.lib()
    L.pa_loglik_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32,
                                    C.c_void_p, C.c_void_p]
    L.pa_loglik_noalloc.restype = C.c_int
    bufs = [_Dev(hip, np.ascontiguousarray(ids, np.int32)), _Dev(hip, np.ascontiguousarray(Z, np.int32)),
            _Dev(hip, np.ascontiguousarray(rates, np.float64)), _Dev(hip, np.full(len(ids), -1.0))]
    st = L.pa_loglik_noalloc(eng._h, bufs[0].p, len(ids), bufs[1].p, bufs[2].p, int(prec), bufs[3].p, None)
    eng._check(st)
    assert hip.hipDeviceSynchronize() == 0
    out = bufs[3].get(np.float64, len(ids))
    for b in bufs:
        b.free()
    return out


def _random_Z(eng, ids, rng):
    Z = np.full((len(ids), eng.N), -1, np.int32)
    for k, i in enumerate(ids):
        d = eng.element(int(i))
        Z[k, d["nodes"]] = rng.integers(0, 3, size=len(d["nodes"]))
        Z[k, d["root"]] = rng.integers(0, 2)
    return Z


def test_noalloc_matches_loglik(require_gpu):
    hip = _hip()
    w = synth.make_workload(2, seed=23, n_elements=40)
    eng = E.Engine.from_workload(w)
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    ids = ids[st == E.PA_ELEM_OK]
    rng = np.random.default_rng(5)
    Z = _random_Z(eng, ids, rng)
    rates = np.stack([rng.uniform(0.1, 0.9, len(ids)), rng.uniform(1.2, 4.0, len(ids))], 1)
    # the arenas reach the device with an ordinary call
    ref64 = eng.loglik(ids, Z, rates, E.PA_FP64)
    ref32 = eng.loglik(ids, Z, rates, E.PA_FP32)
    # two calls in a row, the second smaller (the array keeps the first call's contents)
    for sel in (np.arange(len(ids)), np.arange(0, len(ids), 5)):
        got64 = _noalloc(eng, hip, ids[sel], Z[sel], rates[sel], E.PA_FP64)
        got32 = _noalloc(eng, hip, ids[sel], Z[sel], rates[sel], E.PA_FP32)
        # same kernel and partials; only the block size may reorder the FP64 sum over patterns
        np.testing.assert_allclose(got64, ref64[sel], rtol=1e-12)
        np.testing.assert_allclose(got32, ref32[sel], rtol=1e-12)
    # an id outside the engine's elements scores NaN, the others are unaffected
    bad = np.array([ids[0], 10 ** 6, ids[1]], np.int32)
    got = _noalloc(eng, hip, bad, Z[[0, 0, 1]], rates[[0, 0, 1]], E.PA_FP64)
    assert np.isnan(got[1]) and abs(got[0] - ref64[0]) <= 1e-12 * abs(ref64[0])
    assert abs(got[2] - ref64[1]) <= 1e-12 * abs(ref64[1])
    eng.close()


def test_noalloc_needs_arenas_on_device(require_gpu):
    w = synth.make_workload(2, seed=24, n_elements=4)
    eng = E.Engine.from_workload(w)
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    L = E.lib()
    L.pa_loglik_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32,
                                    C.c_void_p, C.c_void_p]
    L.pa_loglik_noalloc.restype = C.c_int
    # non-null stand-ins: the call must refuse before it touches them
    st = L.pa_loglik_noalloc(eng._h, C.c_void_p(8), 1, C.c_void_p(8), C.c_void_p(8), E.PA_FP64, C.c_void_p(8), None)
    assert st == E.PA_E_STATE
    eng.close()
