"""Continuous request loop: chains served through lanes of concurrent
launches give the same results as one batched run, a chain's result is the
same alone and in a mix of other requests arriving while others run, and
elements retire one by one rather than with their launch."""
import numpy as np
import pytest

from chain_checks import oracle_chains
from oracle import oracle_py as O
from phyloacc_amd import engine as E
from phyloacc_amd import synth

FIELDS = ("max_loglik", "max_m", "status", "final_nrate", "final_crate", "final_grate", "final_lrate",
          "final_lrate2", "accept_n", "accept_c")


def _setup(n, seed=3, config=2, **kw):
    opts = dict(burnin=30, mcmc=90, adapt_freq=25)
    opts.update(kw)
    w = synth.make_workload(config, seed=seed, n_elements=n)
    eng = E.Engine.from_workload(w, E.default_options(**opts))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    return eng, ids[st == E.PA_ELEM_OK]


def _drain(srv, n_expected, max_steps=100000):
    out = {}
    steps = 0
    while len(out) < n_expected:
        srv.step(wait=True)
        for r in srv.poll():
            out[r.ticket] = r
        steps += 1
        assert steps < max_steps
    return out


def _as_tuple(r):
    return tuple(getattr(r.result[k], f) for k in range(3) for f in FIELDS)


def _slot_bytes(eng, P, fp64):
    """HBM scratch of one chain workgroup for elements of up to P patterns
    (padded to 64): partials and exponents (two buffers per internal node),
    the three rate classes' P(t) and log P(t) tables, ambiguity counts."""
    al = lambda x: (x + 255) // 256 * 256
    ppad = (P + 63) // 64 * 64
    I, N, S, rs = eng.N - eng.S, eng.N, eng.S, 8 if fp64 else 4
    return al(2 * I * ppad * 4 * rs) + al(2 * I * ppad * 4) + 2 * al(3 * N * 16 * 8) + al(S * 24 * 4)


def test_lane_count_is_checked():
    eng, _ = _setup(2)
    for bad in (-1, E.PA_SERVE_MAX_LANES + 1):
        with pytest.raises(E.PaError) as ei:
            E.Server(eng, lanes=bad)
        assert ei.value.status == E.PA_E_INPUT and "lanes" in str(ei.value)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,lanes", [(E.PA_FP64, 1), (E.PA_FP64, 3), (E.PA_FP32, 3)])
def test_served_chains_equal_one_batched_run(require_gpu, precision, lanes):
    eng, ids = _setup(16)
    ref = eng.run_chains(ids, precision=precision, block_threads=128)
    srv = E.Server(eng, precision=precision, block_threads=128, max_chains=5, lanes=lanes)
    tickets = srv.submit(ids)
    got = _drain(srv, len(ids))
    assert srv.stats().launches >= 4
    for k, t in enumerate(tickets):
        r = got[t]
        assert r.element == ids[k] and r.kind == E.PA_SERVE_CHAINS
        for ms in range(3):
            for f in FIELDS:
                assert getattr(r.result[ms], f) == ref[f][k, ms], (k, ms, f)
    srv.close()


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [1, 3])
def test_same_result_alone_and_in_a_mix(require_gpu, lanes):
    eng, ids = _setup(24, seed=8)
    x = int(ids[5])
    alone = E.Server(eng, precision=E.PA_FP32, lanes=1)
    t = alone.submit([x])[0]
    ra = _drain(alone, 1)[t]
    alone.close()
    mix = E.Server(eng, precision=E.PA_FP32, max_chains=7, lanes=lanes)
    others = [int(i) for i in ids if i != x]
    tick = list(mix.submit(others[:9]))
    mix.step(wait=False)                      # a launch in flight when x arrives
    tx = mix.submit([x])[0]
    tick += [tx]
    tick += list(mix.submit(others[9:]))
    got = _drain(mix, len(tick))
    got.update({r.ticket: r for r in mix.poll()})
    assert _as_tuple(got[tx]) == _as_tuple(ra)
    st = mix.stats()
    assert st.finished == len(tick) and st.in_flight == 0 and st.waiting == 0
    assert got[tx].arrival_s <= got[tx].admit_s <= got[tx].finish_s
    mix.close()


@pytest.mark.gpu
def test_hbm_budget_limits_chains_per_launch(require_gpu):
    eng, ids = _setup(12, seed=4)
    ref = eng.run_chains(ids, precision=E.PA_FP32, block_threads=128)
    biggest = max(eng.element(int(i))["P"] for i in ids)
    budget = 3.3 * _slot_bytes(eng, biggest, False)
    srv = E.Server(eng, precision=E.PA_FP32, block_threads=128, hbm_budget_bytes=budget)
    tickets = srv.submit(ids)
    got = {}
    peak = 0
    while len(got) < len(ids):
        srv.step(wait=False)
        st = srv.stats()
        assert st.scratch_bytes_in_use <= budget
        peak = max(peak, st.in_flight)
        srv.step(wait=True)
        got.update({r.ticket: r for r in srv.poll()})
    # one scratch slot per workgroup, sized for the launch's longest element:
    # the budget holds 3.3 slots of the longest element overall, so shorter
    # elements share a launch in larger numbers, but never all twelve
    assert 1 <= peak < len(ids)
    assert srv.stats().launches >= 2
    for k, t in enumerate(tickets):
        assert got[t].result[1].max_loglik == ref["max_loglik"][k, 1]
    srv.close()


@pytest.mark.gpu
def test_elements_added_while_serving(require_gpu):
    """Elements encoded after the server started move the device arenas; the
    next launch uploads them and earlier requests are unaffected."""
    w = synth.make_workload(2, seed=12, n_elements=14)
    eng = E.Engine.from_workload(w, E.default_options(burnin=20, mcmc=40, adapt_freq=15))
    ids_a, st_a = eng.add_elements(w.align, w.starts[:8], w.sizes[:8])
    ref_eng = E.Engine.from_workload(w, E.default_options(burnin=20, mcmc=40, adapt_freq=15))
    ref_ids, ref_st = ref_eng.add_elements(w.align, w.starts, w.sizes)
    ref = ref_eng.run_chains(ref_ids[ref_st == E.PA_ELEM_OK], precision=E.PA_FP32, block_threads=128)
    ok_ref = list(ref_ids[ref_st == E.PA_ELEM_OK])
    srv = E.Server(eng, precision=E.PA_FP32, block_threads=128, max_chains=3)
    t_a = list(srv.submit(ids_a[st_a == E.PA_ELEM_OK]))
    got = {}
    for _ in range(2):
        srv.step(wait=True)
        got.update({r.ticket: r for r in srv.poll()})
    ids_b, st_b = eng.add_elements(w.align, w.starts[8:], w.sizes[8:])
    t_b = list(srv.submit(ids_b[st_b == E.PA_ELEM_OK]))
    tick = t_a + t_b
    got.update(_drain(srv, len(tick) - len(got)))
    for t in tick:
        r = got[t]
        k = ok_ref.index(r.element)       # engine ids follow workload order in both engines
        for ms in range(3):
            assert r.result[ms].max_loglik == ref["max_loglik"][k, ms]
            assert r.result[ms].max_m == ref["max_m"][k, ms]
    srv.close()


@pytest.mark.gpu
def test_gene_tree_requests_through_the_loop(require_gpu):
    """Own-tree (gene-tree) log-likelihood requests are served beside chains
    and equal the batched call."""
    gw = synth.make_gt_workload(seed=6, n_elements=20, S=30)
    eng = E.Engine.from_workload(gw, E.default_options(burnin=10, mcmc=30, adapt_freq=15))
    gids, gst = eng.add_elements_with_trees(gw.align, gw.starts, gw.sizes, gw.children())
    L = gw.lengths()
    ref = eng.loglik_lengths(gids, L, E.PA_FP32)
    cids, cst = eng.add_elements(gw.align, gw.starts[:6], gw.sizes[:6])   # species-tree chains alongside
    srv = E.Server(eng, precision=E.PA_FP32)
    tc = list(srv.submit(cids[cst == E.PA_ELEM_OK]))
    tl = list(srv.submit_lengths(gids[:10], L[:10]))
    srv.step(wait=True)
    tl += list(srv.submit_lengths(gids[10:], L[10:]))
    got = _drain(srv, len(tc) + len(tl))
    for k, t in enumerate(tl):
        assert got[t].kind == E.PA_SERVE_LOGLIK
        # same kernel; the batched call may pick another block size, which only
        # reorders the FP64 sum over patterns
        np.testing.assert_allclose(got[t].loglik, ref[k], rtol=1e-12)
    assert all(got[t].kind == E.PA_SERVE_CHAINS for t in tc)
    srv.close()


@pytest.mark.gpu
def test_served_fp64_chains_match_oracle(require_gpu):
    """FP64 chains served while other requests are in flight (requests arrive
    after a launch has started) reproduce the oracle's chains: maximum
    log-likelihood, its state, the final rates and the adaptive-proposal
    counters (src/PhyloAcc-ST/bpp_c.cpp:470-485)."""
    kw = dict(burnin=10, mcmc=30, adapt_freq=12)
    w = synth.make_workload(2, seed=21, n_elements=10)
    eng = E.Engine.from_workload(w, E.default_options(**kw))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    ok = st == E.PA_ELEM_OK
    ids, idx = ids[ok], np.flatnonzero(ok)
    refs = oracle_chains(O.model_from_workload(w), O.default_options(**kw), w, idx)
    srv = E.Server(eng, precision=E.PA_FP64, max_chains=4)
    tick = list(srv.submit(ids[:4]))
    srv.step(wait=False)                       # a launch in flight
    tick += list(srv.submit(ids[4:]))          # arrivals while it runs
    got = _drain(srv, len(tick))
    for k, t in enumerate(tick):
        r = got[t]
        assert r.element == ids[k] and r.status == E.PA_ELEM_OK
        for ms in range(3):
            o = refs[k][ms]
            c = r.result[ms]
            assert abs(c.max_loglik - o["max_loglik"]) <= 1e-9 * abs(o["max_loglik"])
            assert c.max_m == o["max_m"]
            for f, tr in (("final_nrate", "nrate"), ("final_crate", "crate"), ("final_grate", "grate"),
                          ("final_lrate", "lrate"), ("final_lrate2", "l2rate")):
                assert abs(getattr(c, f) - o[tr][-1]) <= 1e-9 * abs(o[tr][-1]), (k, ms, f)
            assert (c.accept_n, c.accept_c) == (o["accept_n"], o["accept_c"])
    srv.close()


@pytest.mark.gpu
def test_lanes_overlap_and_elements_retire_one_by_one(require_gpu):
    """Requests that arrive while a launch runs start in a second lane at once
    (two launches in flight), and a short element of a launch retires before
    its launch's longest element has finished."""
    w = synth.make_workload(3, seed=3, lengths=[60, 70, 80, 9958])
    eng = E.Engine.from_workload(w, E.default_options(burnin=40, mcmc=160, adapt_freq=50))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    ids = ids[st == E.PA_ELEM_OK]
    ref = eng.run_chains(ids, precision=E.PA_FP32, block_threads=128)
    srv = E.Server(eng, precision=E.PA_FP32, block_threads=128, lanes=2)
    t_first = list(srv.submit(ids[[0, 3]]))        # a short and the long element in one launch
    srv.step(wait=False)
    t_second = list(srv.submit(ids[[1, 2]]))       # arrive while that launch runs
    srv.step(wait=False)
    assert srv.stats().lanes_busy == 2 and srv.stats().launches == 2
    got = {}
    while len(got) < 4:
        srv.step(wait=False)
        for r in srv.poll():
            got[r.ticket] = r
    tick = t_first + t_second
    order = [0, 3, 1, 2]
    for t, k in zip(tick, order):
        for ms in range(3):
            for f in FIELDS:
                assert getattr(got[t].result[ms], f) == ref[f][k, ms], (k, ms, f)
    # the three short elements (two of them admitted later) finish before the long one
    assert max(got[t].finish_s for t in (t_first[0], *t_second)) < got[t_first[1]].finish_s
    srv.step(wait=True)                            # the lanes end once their launches have
    st = srv.stats()
    assert st.in_flight == 0 and st.lanes_busy == 0 and st.scratch_bytes_in_use == 0
    srv.close()
