"""Prepared batches: the batched chain launch captured once as a HIP graph
and replayed.  Every replay must equal the eager pa_run_chains result bit for
bit (same streams, same block size), also with eager work and
synchronisations between replays, and a batch refuses to run once the
element arenas have moved."""
import numpy as np
import pytest

from phyloacc_amd import engine as E
from phyloacc_amd import synth

FIELDS = [f for f, _ in E.RESULT_FIELDS]


def _setup(seed=5, n=12):
    w = synth.make_workload(2, seed=seed, n_elements=n)
    eng = E.Engine.from_workload(w, E.default_options(burnin=20, mcmc=60, adapt_freq=25))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    return w, eng, ids[st == E.PA_ELEM_OK]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [E.PA_FP32, E.PA_FP64])
def test_graph_replays_equal_eager_runs(require_gpu, precision):
    w, eng, ids = _setup()
    eager = eng.run_chains(ids, precision=precision, block_threads=128)
    b = E.Batch(eng, ids, precision=precision, block_threads=128, keep_traces=precision == E.PA_FP32)
    rng = np.random.default_rng(1)
    for rep in range(3):
        got = b.run()
        for f in FIELDS:
            np.testing.assert_array_equal(got[f], eager[f], err_msg=f"replay {rep} field {f}")
        assert got["stats"]["algorithmic_bytes"] == eager["stats"]["algorithmic_bytes"]
        assert got["stats"]["n_launches"] == 1 and got["elapsed_ms"] > 0
        # eager work on the same stream between replays
        sub = ids[rng.permutation(len(ids))[:5]]
        again = eng.run_chains(sub, precision=precision, block_threads=128)
        k = [list(ids).index(i) for i in sub]
        np.testing.assert_array_equal(again["max_loglik"], eager["max_loglik"][k])
    b.close()


@pytest.mark.gpu
def test_batch_refuses_moved_arenas(require_gpu):
    w, eng, ids = _setup(seed=6, n=6)
    b = E.Batch(eng, ids)
    b.run()
    eng.add_elements(w.align, w.starts[:2], w.sizes[:2])
    with pytest.raises(E.PaError) as ei:
        b.run()
    assert ei.value.status == E.PA_E_STATE
    eng.close()                      # frees the batch with the engine
    b.close()
