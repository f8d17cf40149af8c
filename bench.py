"""Benchmark: element-MCMC iterations/s of the batched MI355X engine.

Workload (default): 43 species, 50,000 elements with Pareto(1.5) lengths on
[50, 10000] bp, 1-20% of branches accelerated per element (synthetic,
seeded).  One step = the full three-model chain (M0, M1, M2; BURNIN 50 +
MCMC 150 states each by default) of every element in one persistent
launch.  One element-MCMC iteration = one sweep m -> m+1 of one
model chain of one element; value sums the three models.

Also reported (same JSON line):
  roofline     - the chain kernel's counted algorithmic HBM bytes per launch
                 over its HIP-event duration, against 8 TB/s;
  cpu_baseline - the CPU oracle (FP64, per site, OpenMP over elements, all
                 the process's cores) on a length-stratified sample of the
                 same workload;
  plain_gpu    - the plain FP64 path: one launch per element, one pattern
                 per site, on a length-stratified sample;
  fp64_batched - the same workload and launch at FP64 partials;
  reference_chain - 200 + 800-state chains (the reference's defaults, with
                 the adaptive-proposal step at ADAPT_FREQ 500 inside the
                 timed region) against 200-state chains on a length-stratified
                 5,000-element subset: the per-sweep cost does not depend on
                 the chain length, so the headline keeps 200-state chains;
  config2      - the 5,000-element 43-species workload, batched FP32;
  config4      - the continuous serving loop: config-2 elements arriving as a
                 Poisson stream at 2,000/s with 2,000-state chains,
                 throughput and p50/p99 completion latency;
  config5      - 20,000 elements with their own 120-species gene trees:
                 batched log-likelihood evaluations/s, FP32 and FP64 errors
                 against the oracle on a sample, and one sweep of the
                 PhyloAcc-GT gene-tree moves (history, regraft, every node
                 height, pi) over all of them.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stratified(sizes, k, rng):
    """k element indices at evenly spaced length quantiles."""
    order = np.argsort(sizes, kind="stable")
    q = ((np.arange(k) + 0.5) / k * len(order)).astype(int)
    return order[np.clip(q, 0, len(order) - 1)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--elements", type=int, default=None, help="scale the element count down")
    ap.add_argument("--burnin", type=int, default=50)
    ap.add_argument("--mcmc", type=int, default=150)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=48)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-burnin", type=int, default=50)
    ap.add_argument("--cpu-mcmc", type=int, default=550)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--copy-probe", action="store_true", help="run a plain copy kernel of known size and exit")
    ap.add_argument("--no-tune", action="store_true", help="skip the block-size autotuner")
    ap.add_argument("--no-agree", action="store_true", help="skip the FP32-vs-FP64 agreement sample")
    ap.add_argument("--agree-sample", type=int, default=256)
    ap.add_argument("--arrival-rate", type=float, default=2000.0, help="config 4: requests per second")
    ap.add_argument("--chain", type=int, default=None, help="config 4: states per model chain (BURNIN = 1/4)")
    ap.add_argument("--hbm-budget-gib", type=float, default=16.0, help="config 4: scratch budget")
    ap.add_argument("--lanes", type=int, default=0, help="config 4: launches in flight (0 = library default)")
    ap.add_argument("--no-graph", action="store_true", help="time direct launches instead of the captured graph")
    ap.add_argument("--gt-cache", default=None,
                    help="config 5: load the gene-tree workload from this file, or simulate and save it there")
    ap.add_argument("--gt-sweeps", type=int, default=1,
                    help="config 5: timed sweeps of the gene-tree moves (0 skips them)")
    ap.add_argument("--table", action="store_true",
                    help="per-configuration figures (batched FP32/FP64, plain FP64, CPU, agreement) as one JSON line")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the fp64, reference-chain, config 2, 4 and 5 blocks of the default line")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs shard elements across "
                          "processes and are not part of this build", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.copy_probe:
        return copy_probe(args)

    from phyloacc_amd import engine as E
    from phyloacc_amd import synth

    if args.table:
        return table(args, E, synth)
    if args.config == 4:
        print(json.dumps(serve(args, E, synth, args.elements, args.arrival_rate, args.chain, args.lanes)))
        return
    if args.config == 5:
        return gene_trees(args, E, synth)
    extra = args.config == 3 and not args.no_extra
    # the gene-tree workload is simulated in forked workers, before this
    # process touches the GPU
    gt_w = make_gt(synth, args.seed, 20000, args.gt_cache) if extra else None

    t0 = time.time()
    w = synth.make_workload(args.config, seed=args.seed, n_elements=args.elements)
    log(f"workload config{args.config}: {w.C} elements, {int(w.sizes.sum())} sites, S={w.tree.S} "
        f"({time.time() - t0:.1f}s)")
    opts = E.default_options(burnin=args.burnin, mcmc=args.mcmc)
    eng = E.Engine.from_workload(w, opts)
    t1 = time.time()
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    keep = ids[st == E.PA_ELEM_OK]
    log(f"encoded {len(ids)} elements ({len(keep)} kept) in {time.time() - t1:.1f}s")
    prec = E.PA_FP32 if args.precision == "fp32" else E.PA_FP64
    T = args.burnin + args.mcmc
    sweeps_per_step = len(keep) * 3 * (T - 1)

    tuning = None
    if args.block == 0 and not args.no_tune:
        tt = time.time()
        tuning = eng.autotune(keep, precision=prec, states=8)
        log(f"autotune: {tuning} ({time.time() - tt:.1f}s)")
    # the step: one persistent launch over every element, captured once as a
    # HIP graph (clear counters, chain kernel, result copies) and replayed
    batch = None if args.no_graph else E.Batch(eng, keep, precision=prec, block_threads=args.block, keep_traces=True)
    step_fn = batch.run if batch else (lambda: eng.run_chains(keep, precision=prec, block_threads=args.block,
                                                              keep_traces=True))
    for i in range(args.warmup):
        r = step_fn()
        log(f"warmup {i}: {r['elapsed_ms'] / 1e3:.2f}s device")
    ms, stats = [], None
    for i in range(args.steps):
        tw = time.time()
        r = step_fn()
        wall = time.time() - tw
        ms.append(r["elapsed_ms"])
        stats = r["stats"]
        bad = int(np.sum(r["status"] != 0))
        log(f"step {i}: {r['elapsed_ms'] / 1e3:.3f}s device, {wall:.3f}s wall, numeric failures {bad}")
    if batch:
        batch.close()
    step_ms = float(np.mean(ms))
    value = sweeps_per_step / (step_ms / 1e3)
    achieved = stats["algorithmic_bytes"] / (step_ms / 1e3) / 1e9
    traffic, traffic_src = measured_traffic(args, w.C)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": traffic_src,
                "kernel": "chain_kernel", "algorithmic_bytes_per_launch": stats["algorithmic_bytes"],
                "grid": stats["grid"], "block_threads": stats["block_threads"], "lds_bytes": stats["lds_bytes"],
                "launches": stats["n_launches"], "tuning": tuning, "hip_graph": batch is not None}

    rng = np.random.default_rng(0)
    agree = None
    if not args.no_agree and prec == E.PA_FP32:
        agree = fp32_vs_fp64(E, eng, keep, w, args, rng)
        log(f"fp32 vs fp64: {agree}")
    plain = None
    if not args.no_plain:
        # plain FP64 path: per-element launches over an uncompressed copy of a stratified sample
        idx = stratified(w.sizes, 24, rng)
        po = E.default_options(burnin=20, mcmc=80, compress_patterns=0)
        peng = E.Engine.from_workload(w, po)
        pids, pst = peng.add_elements(w.align, w.starts[idx], w.sizes[idx])
        pids = pids[pst == E.PA_ELEM_OK]
        peng.run_chains(pids[:2], precision=E.PA_FP64, mode=E.PA_RUN_PER_ELEMENT)
        rp = peng.run_chains(pids, precision=E.PA_FP64, mode=E.PA_RUN_PER_ELEMENT)
        psweeps = len(pids) * 3 * (po.burnin + po.mcmc - 1)
        mean_pop = float(np.mean(w.sizes))
        mean_smp = float(np.mean(w.sizes[idx][pst == E.PA_ELEM_OK]))
        prate = psweeps / (rp["elapsed_ms"] / 1e3) * (mean_smp / mean_pop)
        plain = {"value": round(prate, 1), "unit": "element-MCMC iterations/s", "precision": "fp64",
                 "launches": int(len(pids)), "sample": f"{len(pids)} length-stratified elements, 100-state chains, "
                 "rate scaled by mean length (sample/population)"}
        peng.close()
        log(f"plain fp64 per-element: {prate:.1f} it/s")

    blocks = {}
    if extra:
        blocks["fp64_batched"] = fp64_batched(E, eng, keep, sweeps_per_step, stats, value)
        blocks["reference_chain"] = reference_chain(E, w, keep, eng, rng)
        blocks["config2"] = config2_block(E, synth, args.seed)
        blocks["config4"] = serve(args, E, synth, None, 2000.0, 2000, args.lanes)
        blocks["config5"] = gene_trees_block(
            E, gt_w, traffic=measured_traffic(args, None, f"config=5 elements={gt_w.C} seed={args.seed} precision=fp32")[0])
        if args.gt_sweeps > 0:
            blocks["config5"]["gene_tree_moves"] = gt_moves_block(E, gt_w, args.gt_sweeps)
        del gt_w
        for k, v in blocks.items():
            log(f"{k}: {v}")

    cpu = None
    if not args.no_cpu:
        from oracle import oracle_py as O
        idx = stratified(w.sizes, args.cpu_sample, rng)
        m = O.model_from_workload(w)
        oo = O.default_options(burnin=args.cpu_burnin, mcmc=args.cpu_mcmc)
        els = [O.Element(m, oo, w.align, w.starts[i], w.sizes[i]) for i in idx]
        els = [e for e in els if not e.filtered]
        host = host_cpu()
        nthr = args.cpu_threads or host["usable"]
        tc = time.time()
        n_sw, _ = O.run_batch(m, oo, els, nthr)
        dt = time.time() - tc
        mean_smp = float(np.mean([e.info()["GG"] for e in els]))
        crate = n_sw / dt * (mean_smp / float(np.mean(w.sizes)))
        cpu = {"value": round(crate, 2), "unit": "element-MCMC iterations/s", "cores": nthr, "kind": "port",
               "nproc": host["nproc"], "cpu_model": host["model"], "core_limit": host["limit"],
               "sample": f"{len(els)} length-stratified elements, {args.cpu_burnin + args.cpu_mcmc}-state chains "
               f"(3 models), {dt:.1f}s; rate scaled by "
               "mean length (sample/population)"}
        log(f"cpu oracle: {crate:.1f} it/s on {nthr} threads ({dt:.1f}s)")

    out = {
        "metric": "element-MCMC iterations/s (batched MI355X engine)",
        "value": round(value, 1),
        "unit": "element-MCMC iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 partials / fp64 sums" if prec == E.PA_FP32 else "fp64",
        "data": "synthetic (seeded simulation of the configuration's shapes)",
        "config": {"workload": f"config{args.config}: S={w.tree.S}, {w.C} elements, lengths "
                   f"{'Pareto(1.5) 50-10000 bp' if args.config == 3 else 'see synth.py'}, "
                   f"BURNIN {args.burnin} + MCMC {args.mcmc}, 3 models",
                   "elements": int(len(keep)), "sites": int(w.sizes.sum()), "parallelism": "1 GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_gpu": plain,
        "speedup_vs_plain_gpu": round(value / plain["value"], 1) if plain else None,
        "fp32_vs_fp64": agree,
    }
    out.update(blocks)
    print(json.dumps(out))


def host_cpu():
    """Cores this process may use (its affinity set, capped by OMP_NUM_THREADS
    when that is set: the host's CPU share), nproc and the CPU model."""
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    limit = os.environ.get("OMP_NUM_THREADS")
    if limit and limit.isdigit() and int(limit) > 0:
        usable = min(usable, int(limit))
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"usable": usable, "nproc": os.cpu_count(), "model": model, "limit": limit}


def make_gt(synth, seed, n, cache=None):
    """The gene-tree workload, simulated in forked workers (so before the
    process touches the GPU), or loaded from `cache`, a file this function
    wrote for the same seed and size."""
    import pickle
    t0 = time.time()
    if cache and os.path.exists(cache):
        with open(cache, "rb") as f:
            w = pickle.load(f)
        assert w.seed == seed and w.C == n, "gene-tree cache holds another workload"
    else:
        w = synth.make_gt_workload(seed=seed, n_elements=n, workers=host_cpu()["usable"])
        if cache:
            with open(cache, "wb") as f:
                pickle.dump(w, f)
    log(f"gene-tree workload: {w.C} elements, S={w.tree.S}, {int(w.sizes.sum())} sites ({time.time() - t0:.1f}s)")
    return w


def fp64_batched(E, eng, keep, sweeps, stats32, value32):
    """The headline workload and launch with FP64 partials (one step)."""
    eng.run_chains(keep[:64], precision=E.PA_FP64)
    r = eng.run_chains(keep, precision=E.PA_FP64, keep_traces=True)
    v = sweeps / (r["elapsed_ms"] / 1e3)
    return {"value": round(v, 1), "unit": "element-MCMC iterations/s", "ms_per_step": round(r["elapsed_ms"], 3),
            "fp32_over_fp64": round(value32 / v, 3), "block_threads": r["stats"]["block_threads"],
            "numeric_failures": int(np.sum(r["status"] != 0))}


def reference_chain(E, w, keep, eng, rng):
    """200 + 800-state chains with ADAPT_FREQ 500 (the reference's defaults)
    against 200-state chains on the same length-stratified 5,000 elements."""
    sizes = np.array([eng.element(int(i))["GG"] for i in keep])
    sub = np.sort(keep[stratified(sizes, min(5000, len(keep)), rng)])
    out = {"elements": int(len(sub))}
    for burnin, mcmc in ((50, 150), (200, 800)):
        opts = E.default_options(burnin=burnin, mcmc=mcmc)     # ADAPT_FREQ 500
        e2 = E.Engine.from_workload(w, opts)
        ids, st = e2.add_elements(w.align, w.starts[sub], w.sizes[sub])
        ids = ids[st == E.PA_ELEM_OK]
        e2.run_chains(ids[:64])
        r = e2.run_chains(ids, keep_traces=True)
        out[f"states_{burnin + mcmc}"] = round(len(ids) * 3 * (burnin + mcmc - 1) / (r["elapsed_ms"] / 1e3), 1)
        e2.close()
    out["ratio_1000_over_200"] = round(out["states_1000"] / out["states_200"], 4)
    out["unit"] = "element-MCMC iterations/s"
    return out


def config2_block(E, synth, seed):
    """The 5,000-element 43-species workload (LogNormal lengths), batched FP32
    and FP64, 200-state chains, at the autotuned launch shape."""
    w = synth.make_workload(2, seed=seed)
    eng = E.Engine.from_workload(w, E.default_options(burnin=50, mcmc=150))
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    ids = ids[st == E.PA_ELEM_OK]
    out = {"elements": int(len(ids)), "unit": "element-MCMC iterations/s",
           "tuning": eng.autotune(ids, precision=E.PA_FP32, states=8)}
    for name, prec in (("fp32", E.PA_FP32), ("fp64", E.PA_FP64)):
        eng.run_chains(ids[:64], precision=prec)
        r = eng.run_chains(ids, precision=prec, keep_traces=True)
        out[name] = round(len(ids) * 3 * 199 / (r["elapsed_ms"] / 1e3), 1)
    eng.close()
    return out


def table(args, E, synth):
    """Per-configuration figures: for configs 1, 2 and 3 the batched FP32 and
    FP64 chain rates over the whole workload, the plain FP64 per-element
    rate and the CPU oracle's rate (all the process's cores) on
    length-stratified samples, and FP32-against-FP64 agreement on a sample;
    for config 5 the gene-tree likelihood rates (batched FP32 and FP64, one
    launch per element in FP64, the oracle on the CPU) and the errors against
    the oracle.  Config 1 runs the reference's default chain (BURNIN 200 +
    MCMC 800); configs 2 and 3 the headline's 200-state chain."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle_py as O
    host = host_cpu()
    nthr = args.cpu_threads or host["usable"]
    gt_w = make_gt(synth, args.seed, 20000)
    rng = np.random.default_rng(0)
    out = {"host": host, "cpu_threads": nthr, "configs": {}}
    for cfg, (burnin, mcmc), cpu_n in ((1, (200, 800), None), (2, (50, 150), 48), (3, (50, 150), 48)):
        t0 = time.time()
        w = synth.make_workload(cfg, seed=args.seed)
        eng = E.Engine.from_workload(w, E.default_options(burnin=burnin, mcmc=mcmc))
        ids, st = eng.add_elements(w.align, w.starts, w.sizes)
        keep = ids[st == E.PA_ELEM_OK]
        T = burnin + mcmc
        sweeps = len(keep) * 3 * (T - 1)
        tune = eng.autotune(keep, precision=E.PA_FP32, states=8)
        row = {"elements": int(len(keep)), "states": T, "tuning": tune}
        for name, prec in (("batched_fp32", E.PA_FP32), ("batched_fp64", E.PA_FP64)):
            eng.run_chains(keep[:64], precision=prec)
            r = eng.run_chains(keep, precision=prec, keep_traces=True)
            row[name] = round(sweeps / (r["elapsed_ms"] / 1e3), 1)
        # agreement: the same chains at both precisions on a stratified sample
        sizes = np.array([eng.element(int(i))["GG"] for i in keep])
        idx = keep[stratified(sizes, min(256, len(keep)), rng)]
        r32 = eng.run_chains(idx, precision=E.PA_FP32, traces=True)
        r64 = eng.run_chains(idx, precision=E.PA_FP64, traces=True)
        b = burnin
        def modes(Z):
            cnt = np.stack([(Z[:, :, b:, :] == z).sum(axis=2) for z in (-1, 0, 1, 2)], axis=-1)
            return cnt.argmax(-1)
        row["fp32_vs_fp64"] = {
            "chains": int(3 * len(idx)),
            "max_rel_err_max_loglik": float(np.max(np.abs(r32["max_loglik"] - r64["max_loglik"]) / np.abs(r64["max_loglik"]))),
            "identical_Z_trajectories": round(float(np.all(r32["Z"] == r64["Z"], axis=(2, 3)).mean()), 4),
            "identical_Z_posterior_modes": round(float(np.all(modes(r32["Z"]) == modes(r64["Z"]), axis=-1).mean()), 4)}
        eng.close()
        # plain FP64: one launch per element over uncompressed sites, 24 stratified elements
        pidx = stratified(w.sizes, min(24, w.C), rng)
        po = E.default_options(burnin=burnin, mcmc=mcmc, compress_patterns=0)
        peng = E.Engine.from_workload(w, po)
        pids, pst = peng.add_elements(w.align, w.starts[pidx], w.sizes[pidx])
        pids = pids[pst == E.PA_ELEM_OK]
        peng.run_chains(pids[:2], precision=E.PA_FP64, mode=E.PA_RUN_PER_ELEMENT)
        rp = peng.run_chains(pids, precision=E.PA_FP64, mode=E.PA_RUN_PER_ELEMENT)
        scale = float(np.mean(w.sizes[pidx][pst == E.PA_ELEM_OK])) / float(np.mean(w.sizes))
        row["plain_fp64"] = round(len(pids) * 3 * (T - 1) / (rp["elapsed_ms"] / 1e3) * scale, 1)
        peng.close()
        # CPU oracle: config 1 in full, the others on a stratified sample
        cidx = np.arange(w.C) if cpu_n is None else stratified(w.sizes, cpu_n, rng)
        m = O.model_from_workload(w)
        oo = O.default_options(burnin=burnin if cpu_n is None else 50, mcmc=mcmc if cpu_n is None else 550)
        els = [O.Element(m, oo, w.align, w.starts[i], w.sizes[i]) for i in cidx]
        els = [e for e in els if not e.filtered]
        tc = time.time()
        n_sw, _ = O.run_batch(m, oo, els, nthr)
        dt = time.time() - tc
        scale = float(np.mean([e.info()["GG"] for e in els])) / float(np.mean(w.sizes))
        row["cpu"] = {"value": round(n_sw / dt * scale, 2), "cores": nthr, "elements": len(els),
                      "states": int(oo.burnin + oo.mcmc), "seconds": round(dt, 1)}
        row["speedup_batched_fp32_vs_plain"] = round(row["batched_fp32"] / row["plain_fp64"], 1)
        row["wall_s"] = round(time.time() - t0, 1)
        out["configs"][str(cfg)] = row
        log(f"config {cfg}: {row}")
    # config 5: gene trees
    w = gt_w
    eng = E.Engine.from_workload(w, E.default_options())
    ids, st = eng.add_elements_with_trees(w.align, w.starts, w.sizes, w.children())
    keep = st == E.PA_ELEM_OK
    ids = ids[keep]
    L = w.lengths()[keep]
    row = {"elements": int(len(ids)), "unit": "element log-likelihoods/s"}
    for name, prec in (("batched_fp32", E.PA_FP32), ("batched_fp64", E.PA_FP64)):
        eng.loglik_lengths(ids, L, prec)
        eng.loglik_lengths(ids, L, prec)
        row[name + "_kernel"] = round(len(ids) / (eng.loglik_stats()["elapsed_ms"] / 1e3), 1)
    full = np.flatnonzero(keep)
    pidx = stratified(w.sizes[keep], 24, rng)
    t1 = time.time()
    for i in pidx:
        eng.loglik_lengths(ids[i:i + 1], L[i:i + 1], E.PA_FP64)
    row["plain_fp64_per_element_calls"] = round(len(pidx) / (time.time() - t1), 1)
    m = O.parse_mod(w.mod_text)[0]
    gp = E.default_options().gap_prop
    smp = stratified(w.sizes[keep], 200, rng)

    def ora(i):
        c = full[i]
        g = w.genes[c]
        return O.gt_loglik(m, w.tree.S, g.children, g.seg_gene, g.seg_species, g.seg_len * w.scale, w.Z[c], w.r1[c],
                           w.r2[c], w.align, w.starts[c], w.sizes[c], gap_prop=gp)
    tc = time.time()
    with ThreadPoolExecutor(nthr) as ex:
        ref = np.array(list(ex.map(ora, smp)))
    dt = time.time() - tc
    row["cpu"] = {"value": round(len(smp) / dt, 1), "cores": nthr, "elements": int(len(smp)), "seconds": round(dt, 2)}
    l32 = eng.loglik_lengths(ids[smp], L[smp], E.PA_FP32)
    l64 = eng.loglik_lengths(ids[smp], L[smp], E.PA_FP64)
    row["fp64_max_rel_err_vs_oracle"] = float(np.max(np.abs(l64 - ref) / np.abs(ref)))
    row["fp32_max_rel_err_vs_oracle"] = float(np.max(np.abs(l32 - ref) / np.abs(ref)))
    eng.close()
    out["configs"]["5"] = row
    log(f"config 5: {row}")
    print(json.dumps(out))


def fp32_vs_fp64(E, eng, keep, w, args, rng):
    """FP32 partials against the FP64 path on a length-stratified sample:
    the same chains (same random streams) at both precisions.  Reports the
    largest relative difference of the per-model maximum log-likelihood, of
    the log-likelihood at a fixed (Z, r), the share of (element, model)
    chains whose whole Z trajectory is identical, and the share whose per-
    branch Z posterior modes (after burn-in) are identical."""
    sizes = np.array([eng.element(int(i))["GG"] for i in keep])
    idx = keep[stratified(sizes, min(args.agree_sample, len(keep)), rng)]
    r32 = eng.run_chains(idx, precision=E.PA_FP32, traces=True)
    r64 = eng.run_chains(idx, precision=E.PA_FP64, traces=True)
    rel = np.abs(r32["max_loglik"] - r64["max_loglik"]) / np.abs(r64["max_loglik"])
    same_traj = np.all(r32["Z"] == r64["Z"], axis=(2, 3))
    b = args.burnin
    def modes(Z):
        cnt = np.stack([(Z[:, :, b:, :] == z).sum(axis=2) for z in (-1, 0, 1, 2)], axis=-1)
        return cnt.argmax(-1)
    same_mode = np.all(modes(r32["Z"]) == modes(r64["Z"]), axis=-1)
    # fixed (Z, r): each element's M2 maximum state, rates at the chain's last values
    Zf = r64["max_Z"][:, 2, :].astype(np.int32)
    rates = np.stack([r64["final_crate"][:, 2], r64["final_nrate"][:, 2]], 1)
    l32 = eng.loglik(idx, Zf, rates, E.PA_FP32)
    l64 = eng.loglik(idx, Zf, rates, E.PA_FP64)
    return {"elements": int(len(idx)), "chains": int(same_traj.size),
            "max_rel_err_max_loglik": float(rel.max()),
            "max_rel_err_loglik_fixed": float(np.max(np.abs(l32 - l64) / np.abs(l64))),
            "identical_Z_trajectories": round(float(same_traj.mean()), 4),
            "identical_Z_posterior_modes": round(float(same_mode.mean()), 4)}


def measured_traffic(args, n_elements, key=None):
    """HBM bytes per launch of this exact workload from a committed counter
    summary under profiles/ (FETCH_SIZE + WRITE_SIZE, scaled by the copy
    probe), or None when no summary covers it."""
    key = key or (f"config={args.config} elements={n_elements} seed={args.seed} burnin={args.burnin} "
                  f"mcmc={args.mcmc} precision={args.precision} block={args.block}")
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None, None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith(".md"):
            continue
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:") and f"(workload: {key})" in line:
                    return float(line.split()[1]), "profiles/" + name
    return None, None


def copy_probe(args):
    """A device-to-device copy of known size (a plain copy kernel), for
    scaling the HBM byte counters of a profiler run: FETCH_SIZE and
    WRITE_SIZE should each read this many bytes per dispatch."""
    import torch
    nbytes = 512 * 2 ** 20
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(max(1, args.steps)):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_per_dispatch": nbytes, "dispatches": max(1, args.steps)}))


def serve(args, E, synth, n_elements, rate, chain, lanes):
    """Continuous serving loop: config-2 elements arrive as a Poisson stream
    at ``rate``/s, each running its three ``chain``-state model chains
    (BURNIN = chain / 4).  The host loop submits arrivals as their time comes
    and turns the server without waiting; every element retires as soon as
    its own chains end.  Reports throughput and per-element completion
    latency (finish - arrival)."""
    w = synth.make_workload(2, seed=args.seed, n_elements=n_elements)
    burnin, mcmc = (args.burnin, args.mcmc) if chain is None else (chain // 4, chain - chain // 4)
    opts = E.default_options(burnin=burnin, mcmc=mcmc)
    eng = E.Engine.from_workload(w, opts)
    ids, st = eng.add_elements(w.align, w.starts, w.sizes)
    ids = ids[st == E.PA_ELEM_OK]
    prec = E.PA_FP32 if args.precision == "fp32" else E.PA_FP64
    T = burnin + mcmc
    rng = np.random.default_rng(args.seed + 1)
    order = rng.permutation(len(ids))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(ids)))
    tuning = None if args.block or args.no_tune else eng.autotune(ids, precision=prec, states=8)
    srv = E.Server(eng, precision=prec, block_threads=args.block, lanes=lanes,
                   hbm_budget_bytes=args.hbm_budget_gib * 2 ** 30)
    # warm the kernel and the library once, outside the timed stream
    srv.submit(ids[order[:4]])
    srv.step(wait=True)
    srv.poll()
    base = srv.stats()
    done = []
    i = 0
    peak_lanes = 0
    t0 = E.serve_clock()
    last_log = t0
    while len(done) < len(ids):
        now = E.serve_clock() - t0
        j = int(np.searchsorted(arrivals, now, side="right"))
        if j > i:
            srv.submit(ids[order[i:j]], t0 + arrivals[i:j])
            i = j
        n = srv.step(wait=False)
        if n:
            done += srv.poll()
        else:
            time.sleep(0.0002)
        if E.serve_clock() - last_log > 10:
            last_log = E.serve_clock()
            s = srv.stats()
            peak_lanes = max(peak_lanes, s.lanes_busy)
            log(f"serve: {len(done)}/{len(ids)} finished, {s.in_flight} in flight, {s.waiting} waiting, "
                f"{s.lanes_busy} lanes busy")
        else:
            peak_lanes = max(peak_lanes, srv.stats().lanes_busy)
    t_end = max(r.finish_s for r in done)
    wall = t_end - t0
    lat = np.array([r.finish_s - r.arrival_s for r in done])
    s = srv.stats()
    sweeps = len(done) * 3 * (T - 1)
    bad = sum(1 for r in done if r.status != E.PA_ELEM_OK)
    srv.close()
    eng.close()
    return {
        "metric": "element-MCMC iterations/s (continuous serving loop)",
        "value": round(sweeps / wall, 1),
        "unit": "element-MCMC iterations/s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 partials / fp64 sums" if prec == E.PA_FP32 else "fp64",
        "data": "synthetic (seeded simulation of the configuration's shapes)",
        "config": {"workload": f"config4: config-2 elements (S={w.tree.S}) arriving as a Poisson stream at "
                   f"{rate:g}/s, BURNIN {burnin} + MCMC {mcmc}, 3 models", "elements": int(len(ids)),
                   "parallelism": "1 GPU"},
        "latency_s": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4),
                      "max": round(float(lat.max()), 4), "mean": round(float(lat.mean()), 4)},
        "serving": {"arrival_rate_per_s": rate, "offered_s": round(float(arrivals[-1]), 3),
                    "launches": int(s.launches - base.launches), "peak_lanes_busy": int(peak_lanes),
                    "tuning": tuning,
                    "device_busy_frac": round((s.device_ms - base.device_ms) / 1e3 / wall, 4),
                    "numeric_failures": bad},
    }


def gene_trees(args, E, synth):
    """Config 5 alone: the gene-tree block as a JSON line of its own."""
    from phyloacc_amd import synth as _s
    w = make_gt(_s, args.seed, args.elements or 20000, args.gt_cache)
    tr, _ = measured_traffic(args, None, f"config=5 elements={w.C} seed={args.seed} precision={args.precision}")
    out = gene_trees_block(E, w, args.steps, args.warmup, args.precision, tr, oracle=not args.no_cpu)
    out.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None})
    if args.gt_sweeps > 0:
        out["gene_tree_moves"] = gt_moves_block(E, w, args.gt_sweeps, args.precision)
    print(json.dumps(out))


def gt_moves_block(E, w, sweeps=1, precision="fp32"):
    """One PhyloAcc-GT sweep of the gene-tree moves over every element, as
    the reference's Gibbs orders them (src/PhyloAcc-GT/bpp_c.cpp:560-680):
    the history draw, one regraft (Sample_tree2) per element, the height move
    (Sample_BranchLen) at every internal gene node, and one sample_pi step;
    every move of every element batched into one launch per move (the
    regraft takes two; the height moves run in a device-resident session
    whose first full pass stores every element's partials, then one path
    pass per move).  Z and the rates stay fixed (the species-branch
    classes of the workload).  Wall clock per phase, one warmup sweep."""
    prec = E.PA_FP32 if precision == "fp32" else E.PA_FP64
    eng = E.Engine.from_workload(w, E.default_options(min_len=1))
    ids, st = eng.add_elements_with_trees(w.align, w.starts, w.sizes, w.children())
    keep = np.flatnonzero(st == E.PA_ELEM_OK)
    ids = ids[keep]
    n, S, N = len(ids), w.tree.S, w.tree.N
    par_sp = np.where(w.tree.parent < 0, N, w.tree.parent)
    sp = E.SpeciesTree.make(S, par_sp, w.sp_height * w.scale, np.full(N, 2 * w.scale))
    ch = np.ascontiguousarray(w.children()[keep])
    par = np.ascontiguousarray(np.stack([np.where(w.genes[c].parent < 0, N, w.genes[c].parent) for c in keep]),
                               dtype=np.int32)
    H = np.ascontiguousarray(np.stack([w.genes[c].height * w.scale for c in keep]))
    Z = np.ascontiguousarray(w.Z[keep])
    rates = np.ascontiguousarray(np.stack([w.r1[keep], w.r2[keep]], 1))
    sizes = w.sizes[keep].astype(np.int32)
    lens = np.ascontiguousarray(np.stack([sizes, sizes // 2], 1), dtype=np.int32)
    piA = np.full(n, 0.25)
    L = np.stack([E.gt_lengths(sp, ch[i], par[i], H[i], Z[i], rates[i, 0], rates[i, 1]) for i in range(n)])
    f0, f1 = eng.loglik_lengths_split(ids, L, lens[:, 0], lens[:, 1], piA=piA, precision=prec)
    ll = np.ascontiguousarray(np.stack([f0, f1], 1))
    lz = np.zeros((n, 2))
    prior = np.array([E.gt_prior(sp, ch[i], par[i], H[i]) for i in range(n)])
    phases = {"history": 0.0, "regraft": 0.0, "heights": 0.0, "pi": 0.0}
    acc = {"regraft": 0, "heights": 0, "pi": 0}
    moves = 0
    kernel_heights = 0.0
    bytes_heights = 0.0
    for m in range(sweeps + 1):
        timed = m > 0
        ind = np.zeros(n, np.int32)
        t0 = time.time()
        eng.gt_sample_history(sp, ids, ch, par, H, Z, rates, lens[:, 0], m, piA=piA, precision=prec, to_host=False)
        t1 = time.time()
        # the (7m + i)-th gene node of element i, the root skipped
        root = np.argmax(par == N, axis=1)
        branch = (m * 7 + np.arange(n)) % (N - 1)
        branch = (branch + (branch >= root)).astype(np.int32)
        a = eng.gt_sample_tree(sp, ids, ch, par, H, Z, rates, branch, ind, lens, None, ll, lz, m, 0, piA=piA,
                               precision=prec)
        t2 = time.time()
        prior = np.array([E.gt_prior(sp, ch[i], par[i], H[i]) for i in range(n)])
        t2b = time.time()
        # the trees stay on the device for the whole pass over the nodes,
        # every move queued at once (pa_gt_session_sweep)
        ses = E.GtSession(eng, sp, ids, ch, par, H, Z, rates, ll, lz, prior, piA=piA)
        acc_h, kms = ses.sweep(np.arange(S, N, dtype=np.int32), ind, lens, m, 1, precision=prec)
        ah = int(acc_h.sum())
        kbytes = eng.loglik_stats()["algorithmic_bytes"]
        H[:], ll[:], lz[:], prior[:] = ses.read()
        ses.close()
        t3 = time.time()
        L = np.stack([E.gt_lengths(sp, ch[i], par[i], H[i], Z[i], rates[i, 0], rates[i, 1]) for i in range(n)])
        t3b = time.time()
        piA, l_pi, ap = eng.sample_pi(ids, L, m, piA, ll[:, 0], precision=prec)
        ll[:, 0] = l_pi
        t4 = time.time()
        if timed:
            phases["history"] += t1 - t0
            phases["regraft"] += t2 - t1
            phases["heights"] += t3 - t2b
            kernel_heights += kms
            bytes_heights += kbytes
            phases["pi"] += t4 - t3b
            acc["regraft"] += int(a.sum())
            acc["heights"] += ah
            acc["pi"] += int(ap.sum())
            moves += n * (1 + (N - S) + 1)
    eng.close()
    tot = sum(phases.values())
    return {"metric": "gene-tree MCMC moves/s (history, regraft, every node height, pi; batched over elements)",
            "value": round(moves / tot, 1), "unit": "element moves/s", "sweeps": sweeps, "elements": n,
            "ms_per_sweep": round(tot / sweeps * 1e3, 1),
            "phase_ms": {k: round(v / sweeps * 1e3, 1) for k, v in phases.items()},
            "heights_kernel_ms": round(kernel_heights / sweeps, 1),
            # the path passes' counted algorithmic bytes over their HIP-event time
            "heights_kernel_roofline": {"bound": "hbm", "achieved": round(bytes_heights / (kernel_heights * 1e-3) / 1e9, 1),
                                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                        "frac": round(bytes_heights / (kernel_heights * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                        "bytes_per_sweep": bytes_heights / sweeps},
            # likelihood launches: history, two for the regraft, the session's
            # first full pass, one path pass per height move, the pi step
            "launches_per_sweep": 1 + 2 + 1 + (N - S) + 1,
            "accept_rate": {k: round(v / (n * sweeps * ((N - S) if k == "heights" else 1)), 4)
                            for k, v in acc.items()},
            "precision": precision, "heights_path": "device-resident session, one sweep call (pa_gt_session_sweep)",
            "data": "synthetic gene trees of the workload; Z and rates fixed"}


def gene_trees_block(E, w, steps=3, warmup=1, precision="fp32", traffic=None, oracle=True):
    """Uneven-tree batched pruning: every element carries its own
    multispecies-coalescent gene tree (120 species).  A step evaluates the
    log-likelihood of every element at its gene-branch lengths in one launch.
    Reports element log-likelihood evaluations per second at both
    precisions, and the FP32 and FP64 errors against the oracle's gene-tree
    pruning on a sample."""
    from oracle import oracle_py as O
    opts = E.default_options()
    eng = E.Engine.from_workload(w, opts)
    ids, st = eng.add_elements_with_trees(w.align, w.starts, w.sizes, w.children())
    keep = st == E.PA_ELEM_OK
    ids = ids[keep]
    L = w.lengths()[keep]
    rates = {}
    for name, prec in (("fp32", E.PA_FP32), ("fp64", E.PA_FP64)):
        for _ in range(warmup):
            eng.loglik_lengths(ids, L, prec)
        times = []
        for _ in range(steps):
            tw = time.time()
            eng.loglik_lengths(ids, L, prec)
            times.append(time.time() - tw)
        rates[name] = (len(ids) / float(np.mean(times)), float(np.mean(times)))
        if name == precision:
            ks = eng.loglik_stats()     # the last timed call's kernel
            ach = ks["algorithmic_bytes"] / (ks["elapsed_ms"] / 1e3) / 1e9
            roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": traffic, "kernel": "loglik_kernel",
                    "algorithmic_bytes_per_launch": ks["algorithmic_bytes"],
                    "kernel_ms": round(ks["elapsed_ms"], 3), "grid": ks["grid"], "block_threads": ks["block_threads"],
                    "lds_bytes": ks["lds_bytes"]}
    vs = None
    if oracle:
        rng = np.random.default_rng(0)
        smp = rng.choice(len(ids), size=min(48, len(ids)), replace=False)
        m = O.parse_mod(w.mod_text)[0]
        full = np.flatnonzero(keep)
        ref = np.array([O.gt_loglik(m, w.tree.S, w.genes[full[i]].children, w.genes[full[i]].seg_gene,
                                    w.genes[full[i]].seg_species, w.genes[full[i]].seg_len * w.scale, w.Z[full[i]],
                                    w.r1[full[i]], w.r2[full[i]], w.align, w.starts[full[i]], w.sizes[full[i]],
                                    gap_prop=opts.gap_prop) for i in smp])
        l32 = eng.loglik_lengths(ids[smp], L[smp], E.PA_FP32)
        l64 = eng.loglik_lengths(ids[smp], L[smp], E.PA_FP64)
        vs = {"sample": int(len(smp)), "fp64_max_rel_err": float(np.max(np.abs(l64 - ref) / np.abs(ref))),
              "fp32_max_rel_err": float(np.max(np.abs(l32 - ref) / np.abs(ref)))}
    eng.close()
    main = rates[precision]
    return {
        "metric": "gene-tree element log-likelihood evaluations/s (uneven per-element trees, batched)",
        "value": round(main[0], 1), "unit": "element log-likelihoods/s", "ms_per_step": round(main[1] * 1e3, 3),
        "fp64_value": round(rates["fp64"][0], 1),
        "dtype": "fp32 partials / fp64 sums" if precision == "fp32" else "fp64",
        "data": "synthetic (seeded multispecies-coalescent gene trees)",
        "config": {"workload": f"config5: S={w.tree.S} species tree in coalescent units, {len(ids)} elements with "
                   "their own gene trees, lengths LogNormal(ln 300, 0.6)", "elements": int(len(ids)),
                   "sites": int(w.sizes[keep].sum()), "parallelism": "1 GPU",
                   "timing": "wall clock around the call, host-to-device transfer of the lengths included"},
        "vs_oracle": vs,
        "roofline": roof,
    }


if __name__ == "__main__":
    main()
