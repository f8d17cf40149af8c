"""Shared checks of engine chains against the oracle's chains.

The oracle (oracle/pa_oracle.c) restates the reference's per-element MCMC in
FP64 log space (src/PhyloAcc-ST/bpp_c.cpp:346-519) and draws from the same
counter-based random streams as the engine, so a chain of the engine and the
oracle's chain of the same element can be compared state by state.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from oracle import oracle_py as O


def oracle_chains(m, oo, w, idx):
    """The oracle's three model chains (with traces) of elements idx of
    workload w, run in parallel threads (the oracle releases the GIL)."""
    def one(i):
        return O.Element(m, oo, w.align, w.starts[i], w.sizes[i]).run(oo)
    nt = max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(nt) as ex:
        return list(ex.map(one, [int(i) for i in idx]))


def posterior_modes(Z, burnin):
    """Per-branch mode of Z over the states after burn-in; Z: [..., T, N]."""
    post = Z[..., burnin:, :]
    cnt = np.stack([(post == z).sum(axis=-2) for z in (-1, 0, 1, 2)], axis=-1)
    return cnt.argmax(-1) - 1


def compare(res, refs, burnin, rtol):
    """Engine chains res (run_chains output with traces) against oracle chains
    refs (one list of three model dicts per element).  Returns the share of
    chains with identical Z trajectories, the share with identical per-branch
    Z posterior modes and the largest relative error of the per-model maximum
    log-likelihood and of the log-likelihood trace (over chains whose
    trajectories are identical).  rtol bounds the relative errors."""
    n = len(refs)
    same_traj = np.zeros((n, 3), bool)
    same_mode = np.zeros((n, 3), bool)
    err_max, err_trace = 0.0, 0.0
    for k in range(n):
        for ms in range(3):
            r = refs[k][ms]
            same_traj[k, ms] = np.array_equal(res["Z"][k, ms], r["Z"])
            same_mode[k, ms] = np.array_equal(posterior_modes(res["Z"][k, ms], burnin), posterior_modes(r["Z"], burnin))
            err_max = max(err_max, abs(res["max_loglik"][k, ms] - r["max_loglik"]) / abs(r["max_loglik"]))
            if same_traj[k, ms]:
                err_trace = max(err_trace, float(np.max(np.abs(res["loglik"][k, ms] - r["loglik"]) / np.abs(r["loglik"]))))
    return dict(identical_trajectories=float(same_traj.mean()), identical_modes=float(same_mode.mean()),
                max_rel_err_max_loglik=err_max, max_rel_err_loglik_trace=err_trace, rtol=rtol,
                all_modes=bool(same_mode.all()))
