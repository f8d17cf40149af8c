"""Launch shape of batched chain launches: set / get, the autotuner over
block sizes and pattern-loop shapes, and that a shape changes no result."""
import numpy as np
import pytest

from phyloacc_amd import engine as E
from phyloacc_amd import synth

FIELDS = [f for f, _ in E.RESULT_FIELDS]


def _engine(n=6, seed=5, config=2):
    w = synth.make_workload(config, seed=seed, n_elements=n)
    eng = E.Engine.from_workload(w, E.default_options(burnin=10, mcmc=30, adapt_freq=15))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    return eng, ids[st == E.PA_ELEM_OK]


def test_set_and_get_tuning():
    eng, _ = _engine(2)
    t = eng.tuning()
    assert t["block_threads"] == 0 and t["shape"] == 0            # default: by mean pattern count
    eng.set_tuning(256, 2)
    t = eng.tuning()
    assert t["block_threads"] == 256 and t["shape"] == 2 and (t["round_nodes"], t["chunk_nodes"]) == (2, 8)
    eng.set_tuning(0, 0)
    assert eng.tuning()["block_threads"] == 0


def test_invalid_tuning_is_refused():
    eng, _ = _engine(2)
    for bad in ((96, 0), (128, -1), (128, 3)):
        with pytest.raises(E.PaError) as ei:
            eng.set_tuning(*bad)
        assert ei.value.status == E.PA_E_INPUT


@pytest.mark.gpu
def test_autotuned_shape_is_used_and_changes_nothing_else(require_gpu):
    eng, ids = _engine(24)
    t = eng.autotune(ids, precision=E.PA_FP64, states=8)
    bs = t["block_threads"]
    assert bs in (64, 128, 256) and 0 <= t["shape"] < 3 and t["rate"] > 0
    got = eng.run_chains(ids, precision=E.PA_FP64)
    assert got["stats"]["block_threads"] == bs
    ref = eng.run_chains(ids, precision=E.PA_FP64, block_threads=bs)
    for f in FIELDS:
        np.testing.assert_array_equal(got[f], ref[f])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [E.PA_FP32, E.PA_FP64])
def test_every_shape_gives_the_same_chains(require_gpu, precision):
    eng, ids = _engine(16, seed=7)
    outs = []
    for sh in range(3):
        eng.set_tuning(128, sh)
        outs.append(eng.run_chains(ids, precision=precision, traces=True))
    for o in outs[1:]:
        for f in FIELDS + ["Z", "loglik", "nrate"]:
            np.testing.assert_array_equal(o[f], outs[0][f], err_msg=f)
