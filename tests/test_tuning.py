"""Block size of batched chain launches: set / get, the autotuner, and the
switched-off split (two block sizes) configuration."""
import numpy as np
import pytest

from phyloacc_amd import engine as E
from phyloacc_amd import synth


def _engine(n=6, seed=5, config=2):
    w = synth.make_workload(config, seed=seed, n_elements=n)
    eng = E.Engine.from_workload(w, E.default_options(burnin=10, mcmc=30, adapt_freq=15))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    return eng, ids[st == E.PA_ELEM_OK]


def test_set_and_get_one_block_size():
    eng, _ = _engine(2)
    assert eng.tuning()["small_threads"] == 0            # default: by mean pattern count
    eng.set_tuning(256, 256, 0)
    t = eng.tuning()
    assert t["small_threads"] == 256 and t["long_threads"] == 256
    eng.set_tuning(0, 0, 0)
    assert eng.tuning()["small_threads"] == 0


def test_split_block_sizes_are_switched_off():
    eng, _ = _engine(2)
    with pytest.raises(E.PaError) as ei:
        eng.set_tuning(64, 256, 256)
    assert ei.value.status == E.PA_E_STATE and "switched off" in str(ei.value)
    with pytest.raises(E.PaError) as ei:
        eng.set_tuning(96, 96, 0)
    assert ei.value.status == E.PA_E_INPUT


@pytest.mark.gpu
def test_autotuned_block_size_is_used_and_changes_nothing_else(require_gpu):
    eng, ids = _engine(24)
    t = eng.autotune(ids, precision=E.PA_FP64, states=8)
    bs = t["small_threads"]
    assert bs in (64, 128, 256) and t["long_threads"] == bs and t["rate"] > 0
    got = eng.run_chains(ids, precision=E.PA_FP64)
    assert got["stats"]["block_threads"] == bs
    ref = eng.run_chains(ids, precision=E.PA_FP64, block_threads=bs)
    for f in ("max_loglik", "max_m", "accept_n", "accept_c", "final_nrate", "final_crate"):
        np.testing.assert_array_equal(got[f], ref[f])
