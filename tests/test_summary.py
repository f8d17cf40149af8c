"""Eval2 / Output_init in the engine's host code (pa_chain_summary) against
the oracle's restatement, fed the same traces (the oracle's own chain)."""
import numpy as np

from oracle import oracle_py as O
from phyloacc_amd import engine as E
from phyloacc_amd import synth


def test_chain_summary_matches_oracle():
    kw = dict(burnin=30, mcmc=120, adapt_freq=40)
    w = synth.make_workload(1, seed=17, n_elements=6)
    eng = E.Engine.from_workload(w, E.default_options(**kw))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    m = O.model_from_workload(w)
    oo = O.default_options(**kw)
    checked = 0
    for k, i in enumerate(ids):
        ref = O.Element(m, oo, w.align, w.starts[i], w.sizes[i]).run(oo)
        for ms in range(3):
            r = ref[ms]
            tr = dict(full=r["full"], nrate=r["nrate"], crate=r["crate"], grate=r["grate"], lrate=r["lrate"],
                      l2rate=r["l2rate"], Z=r["Z"], max_Z=r["max_Z"])
            # medians and Z frequencies never need the device
            s = eng.chain_summary(i, ms, tr, want_vb=False)
            np.testing.assert_array_equal(s["med"], r["med"])
            np.testing.assert_allclose(s["postZ"], r["postZ"], rtol=0, atol=0)
            # Eval2's main branch is host-only; its fallback re-prunes on the device
            try:
                s2 = eng.chain_summary(i, ms, tr, want_vb=True)
            except E.PaError as exc:
                assert exc.status == E.PA_E_DEVICE
                continue
            assert abs(s2["vb"] - r["vb"]) <= 1e-9 * abs(r["vb"]), (s2["vb"], r["vb"])
            checked += 1
    assert checked > 0
