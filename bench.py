#!/usr/bin/env python3
"""Throughput of the MI355X population step engine on the C2 workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (default): the OoA-style model on a 40,962-cell icosahedral grid,
100,000 initial agents, seed 42; W untimed steps, then K timed steps
(W + K = 2,000 = the full C2 run by default).  A step is one full population
step (initialize, fused actions, births, per-cell recount) on the device.

value = sum over the timed steps of the live count at step start / wall time
of the timed steps (inputs resident in HBM, no host sync inside the region).

Per-kernel device time comes from HIP events attached to each dispatch on the
handle's stream in a second, bit-identical run of the same steps (the
simulation is deterministic); roofline.achieved = algorithmic bytes per launch
(qhg4_amd/csrc/bytes_model.h) / average launch duration of the dominant kernel;
roofline.traffic = its measured HBM bytes per launch (rocprofv3 counters).

cpu_baseline: the CPU oracle (the reference algorithm restated in C with the
same Philox draws, OpenMP) on the box's host cores, timed on the same steps,
bounded to about 20 s.

The same line carries two more objects:
  ensemble  C4: 256 C2-style replicas (1,000 steps each) run one at a time
            through the plain path and as members of one batch; value is
            the batched agent-steps/s, speedup_vs_plain the ratio of the two
            wall times, identical_to_plain the per-replica comparison.
  serving   C5: a Poisson(20/s) stream of 2,000 C2-style requests
            (--serve-requests) served by continuous batching under a 250 GB
            budget; p50/p99 completion latency, aggregate agent-steps/s over
            first arrival .. last completion, and a sample run alone.
and two C2 runs beside the timed window: c2_full_run, the whole 2,000-step
run timed as one region, and c2_stated_patch, the run from the initial state
as stated (100,000 agents in 200 cells), under which the population dies out.
Ensemble and serving results are compared for identity through full-state
digests (every live agent's fields and genome).
--workload c4 / c5 print those measurements as lines of their own.
--workload c2a times the C2 model with genetics and AssortativePairHamming
pairing (k_assort_pair in kernels); --workload c2nav the C2 model with
Navigate's sea-way jumps and bridges.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0

# Measured HBM bytes per algorithmic byte of each step kernel: rocprofv3
# FETCH_SIZE + WRITE_SIZE per dispatch on the C2 workload (each counter
# scaled by what it reads for a 256 MiB device copy: FETCH x2, WRITE x1)
# divided by the kernel's bytes_model.h bytes per dispatch over the same
# steps (profiles/step_kernels_c2_counters.md).  roofline.traffic is this
# ratio times the timed window's own algorithmic bytes per launch.
TRAFFIC_PER_ALGO_BYTE = {"k_cell_act": 1.72, "k_cell_scatter": 1.82, "k_cell_scan": 21.0}
METRIC = "agent-steps/sec at 1 MI355X + %HBM peak; ensemble vs plain speedup, p50/p99"


def bytes_model():
    path = os.path.join(ROOT, "qhg4_amd", "csrc", "bytes_model.h")
    out = {}
    for ln in open(path):
        m = re.match(r"#define\s+QHG_B_(\w+)\s+(\d+)", ln)
        if m:
            out[m.group(1)] = int(m.group(2))
    return out


def kernel_bytes(bm, d, grid, ops, gwords):
    """Algorithmic bytes of each step kernel summed over the measured steps.
    d: deltas of live agent-steps, occupied cell-steps, births, deaths."""
    from qhg4_amd import abi
    conn = grid["connectivity"]
    live, occ, births, deaths, steps = d["live"], d["occupied"], d["births"], d["deaths"], d["steps"]
    per_live = bm["ACT_LIVE_BASE"]
    if abi.ACT_RANDOMPAIR in ops:
        per_live += bm["ACT_LIVE_PAIR"]
    if abi.ACT_GETOLD in ops:
        per_live += bm["ACT_LIVE_GETOLD"]
    if abi.ACT_FERTILITY in ops:
        per_live += bm["ACT_LIVE_FERT"]
    per_occ = bm["ACT_OCC"] + conn * bm["ACT_OCC_PER_CONN"] + bm["ACT_OCC_W_EXTRA"]
    act = occ * per_occ + live * per_live + births * bm["ACT_BIRTH"]
    scan = occ * bm["SCAN_OCC"]
    survivors = live - deaths
    per_surv = bm["SCAT_SURVIVOR"] + (bm["SCAT_SURVIVOR_GEN"] if gwords else 0)
    per_birth = bm["SCAT_BIRTH"] + (bm["SCAT_BIRTH_GEN"] if gwords else 0)
    scat = occ * bm["SCAT_OCC"] + live * bm["SCAT_LIVE"] + survivors * per_surv + births * per_birth
    out = {"k_cell_act": act, "k_cell_scan": scan, "k_cell_scatter": scat}
    if gwords:
        out["k_genome"] = births * bm["SCAT_GENOME_WORDS"] * gwords * 8
    if abi.ACT_ASSORTPAIR in ops and gwords:
        # words per strand position the distance compares (the block count
        # read as a genome size, as AssortativePairHam passes it)
        D = ((gwords // 2) + 31) // 32
        out["k_assort_pair"] = live * bm["ASSORT_LIVE"] + d["fertile"] * (bm["ASSORT_FERTILE"] + 2 * D * 8)
    return out


def make_workload(name, seed):
    from qhg4_amd import scenarios
    if name == "c2":
        sc = scenarios.c2(seed=seed)
        desc = "C2: OoA model (NPPCapacity, GetOld, WeightedMove, MultiEvaluator, Fertility, " \
               "RandomPair, VerhulstVarK) on a 40,962-cell icosahedral grid, 100,000 initial agents"
    elif name == "c1":
        sc = scenarios.c1(seed=seed)
        desc = "C1: plain model on a 128x128 lattice, 10,000 initial agents"
    elif name == "c2a":
        sc = scenarios.c2(seed=seed, genetics=True, assort=(0.9, 1))
        desc = "C2 with 2x256-locus 2-bit genetics and AssortativePairHamming (cutoff 0.9, permuted) " \
               "in place of RandomPair (OoANavGen2bitAssHamPop's pairing), 100,000 initial agents"
    elif name == "c2nav":
        sc = scenarios.c2nav(seed=seed, level=6, n_agents=100000, patch_cells=4000, prob0=0.15, genetics=False)
        desc = "C2 model plus Navigate (OoANavGen2bitPop's sea-ways: ports on the coast with up to 6 " \
               "destinations each, bridges, 3% ice) on a 40,962-cell icosahedral grid, 100,000 initial " \
               "agents on a coast"
    elif name == "c3":
        sc = scenarios.c3(seed=seed)
        desc = "C3: C2 model + 2x256-locus 2-bit genetics on a 655,362-cell icosahedral grid, 2M agents"
    else:
        raise SystemExit("unknown workload %s" % name)
    return sc, desc


C4_DESC = ("C4: ensemble of C2-style replicas (grids of 2,562-163,842 cells, 10k-2M initial agents "
           "log-uniform, per-replica move/birth/death/K-scale parameters), each run alone through the "
           "plain path vs all as members of one batch")


def c4_leg(seed, n, steps, graph=True, tune=True, kernel_timing=True):
    """The ensemble measurement: the same n replicas run one at a time
    through the plain path and all together as members of one batch; the
    per-replica results must be identical."""
    from qhg4_amd import scenarios, ensemble

    def make(r):
        return scenarios.c4_replica(r, base_seed=seed, steps=steps)

    print("bench: C4 %d replicas x %d steps, plain path one at a time" % (n, steps), file=sys.stderr, flush=True)
    t_plain, plain = ensemble.run_plain(
        make, n, steps, log=lambda r, t: print("bench: plain %d done, %.2f s" % (r, t), file=sys.stderr, flush=True))
    print("bench: plain %.2f s; batched" % t_plain, file=sys.stderr, flush=True)
    run = ensemble.BatchedRun(make, n, graph=graph)
    dev_bytes = run.device_bytes()
    t_batch, tuning = run.run(steps, tune_steps=2 if tune else 0, timing_steps=5 if kernel_timing else 0)
    batched = run.summaries()
    identical = batched == plain
    agent_steps = sum(s["agent_steps"] for s in batched)
    print("bench: batched %.2f s, identical=%s" % (t_batch, identical), file=sys.stderr, flush=True)
    # per-kernel device time over the steps timed per dispatch right after
    # tuning, while the members hold their early (largest) populations
    roof, kernels = None, {}
    if run.timed:
        ks = run.timed["steps"]
        bm = bytes_model()
        tot = {}
        for p, a, b in zip(run.pops, run.timed["before"], run.timed["after"]):
            prm = p.params
            ops = {prm.action_kind[i] for i in range(prm.num_actions)}
            dd = delta(b, a, ks)
            for k, v in kernel_bytes(bm, dd, p.grid, ops, 0).items():
                kk = k.replace("k_cell_", "k_batch_")
                tot[kk] = tot.get(kk, 0) + v
            tot["k_batch_scan_sums"] = tot.get("k_batch_scan_sums", 0) + dd["occupied"] * bm["SCAN_SUMS_OCC"]
        kernels, roof = kernel_table(tot, run.timed["times"])
    run.close()
    return {"workload": C4_DESC, "replicas": n, "steps_per_replica": steps, "base_seed": seed,
            "initial_agents_total": int(sum(scenarios.c4_draw(r, seed)["n0"] for r in range(n))),
            "agent_steps_total": int(agent_steps), "device_bytes_resident": int(dev_bytes),
            "value": round(agent_steps / t_batch, 1), "plain_value": round(agent_steps / t_plain, 1),
            "batched_seconds": round(t_batch, 3), "plain_seconds": round(t_plain, 3),
            "speedup_vs_plain": round(t_plain / t_batch, 3), "identical_to_plain": identical, "graph": graph,
            "tuning": tuning, "kernels_early": kernels, "roofline_early": roof,
            "make": make}


def run_c4(args):
    e = c4_leg(args.seed, args.replicas, args.steps, graph=not args.no_graph, tune=not args.no_tune,
               kernel_timing=not args.no_kernel_timing)
    make = e.pop("make")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(make(0), 0, args.steps, args.cpu_seconds)
        cpu["sample"] = "replica 0 of the ensemble: " + cpu["sample"]
    out = base_line(args, e["value"], e["batched_seconds"] * 1e3 / args.steps, 0)
    out["config"] = {k: e[k] for k in ("workload", "replicas", "steps_per_replica", "base_seed",
                                        "initial_agents_total", "agent_steps_total", "device_bytes_resident")}
    out["ensemble"] = {k: e[k] for k in ("plain_value", "plain_seconds", "batched_seconds", "speedup_vs_plain",
                                          "identical_to_plain", "graph", "tuning")}
    out["roofline"] = e["roofline_early"]
    out["cpu_baseline"] = cpu
    out["kernels"] = e["kernels_early"]
    return out


C5_DESC = ("C5: served stream of C2-style scenario requests (Poisson arrivals at 20/s, grids of 2,562-655,362 "
           "cells, 10k-5M initial agents log-uniform, horizons 200-5,000 steps, per-request parameters), "
           "continuous batching under an HBM budget")


def c5_leg(seed, nreq, budget_gb=250.0, chunk=16, plain_sample=16, graph=True):
    """The serving measurement: nreq requests arrive in real time and are
    served by continuous batching; a sample of them is also run alone
    through the plain path, for the plain rate and to check that their
    results are the same alone and in the mix."""
    import numpy as np
    from qhg4_amd import scenarios, ensemble
    from qhg4_amd.server import ScenarioServer

    reqs = scenarios.c5_requests(nreq, base_seed=seed)
    for lv in sorted({r["level"] for r in reqs}):
        scenarios.c4_world(lv, seed)          # the server's grids and fields, built before it starts

    def make(req):
        return scenarios.c5_scenario(req, world_seed=seed)

    def log(msg):
        print("bench: " + msg, file=sys.stderr, flush=True)

    log("C5 %d requests, budget %.0f GB" % (len(reqs), budget_gb))
    srv = ScenarioServer(make, budget_bytes=budget_gb * 1e9, chunk=chunk, graph=graph)
    results, rep = srv.serve(reqs, log=log)
    log("served %d requests in %.1f s, p50 %.2f s p99 %.2f s" % (rep["requests"], rep["seconds"], rep["p50_s"],
                                                                 rep["p99_s"]))
    by_id = {r["id"]: r for r in results}
    sample = [reqs[i] for i in np.linspace(0, len(reqs) - 1, min(plain_sample, len(reqs))).astype(int)]
    t_plain, steps_plain, same = 0.0, 0, True
    for q in sample:
        sec, summ = ensemble.run_plain(lambda _r, q=q: make(q), 1, q["horizon"])
        t_plain += sec
        steps_plain += summ[0]["agent_steps"]
        same = same and summ[0] == by_id[q["id"]]["summary"]
    log("plain sample of %d: %.2f s, identical=%s" % (len(sample), t_plain, same))
    plain_value = steps_plain / t_plain if t_plain > 0 else None
    return {"workload": C5_DESC, "requests": len(reqs), "budget_gb": budget_gb, "chunk_steps": chunk,
            "base_seed": seed, "value": round(rep["value"], 1),
            "p50_latency_s": round(rep["p50_s"], 3), "p99_latency_s": round(rep["p99_s"], 3),
            "mean_latency_s": round(rep["mean_s"], 3), "seconds": rep["seconds"],
            "arrival_window_s": round(reqs[-1]["arrival"], 3), "agent_steps": rep["agent_steps"],
            "peak_members": rep["peak_members"], "peak_device_bytes": rep["peak_device_bytes"],
            "tunings": rep["tunings"], "plain_sample": len(sample),
            "plain_value": round(plain_value, 1) if plain_value else None,
            "speedup_vs_plain": round(rep["value"] / plain_value, 3) if plain_value else None,
            "sample_identical_alone_and_in_mix": same, "make": make, "first": reqs[0]}


def run_c5(args):
    e = c5_leg(args.seed, args.requests, args.budget_gb, args.chunk, args.plain_sample, graph=not args.no_graph)
    make, first = e.pop("make"), e.pop("first")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(make(first), 0, first["horizon"], args.cpu_seconds)
        cpu["sample"] = "request 0 of the stream: " + cpu["sample"]
    out = base_line(args, e["value"], float("nan"), 0)
    out["steps"] = None
    out["ms_per_step"] = None
    out["config"] = {k: e[k] for k in ("workload", "requests", "budget_gb", "chunk_steps", "base_seed")}
    out["serving"] = {k: v for k, v in e.items() if k not in out["config"] and k != "value"}
    out["roofline"] = None
    out["cpu_baseline"] = cpu
    return out


def kernel_table(total_bytes, times):
    kernels = {}
    for name, bts in total_bytes.items():
        avg_ms, n = times.get(name, (0.0, 0))
        if not n:
            continue
        per_launch = bts / n
        gbps = per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        kernels[name] = {"avg_us": round(avg_ms * 1e3, 3), "launches": n,
                         "bytes_per_launch": int(per_launch), "GBps": round(gbps, 1)}
    if not kernels:
        return kernels, None
    dom = max(kernels, key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
    kd = kernels[dom]
    roof = {"bound": "hbm", "achieved": kd["GBps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(kd["GBps"] / HBM_PEAK_GBPS, 5), "traffic": None, "kernel": dom,
            "bytes_per_launch": kd["bytes_per_launch"], "avg_launch_us": kd["avg_us"]}
    return kernels, roof


def base_line(args, value, ms_per_step, warmup):
    return {"metric": METRIC, "value": round(value, 1), "unit": "agent-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": warmup, "ms_per_step": round(ms_per_step, 5),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic"}


def delta(s1, s0, steps):
    return {"live": s1["agent_steps"] - s0["agent_steps"],
            "occupied": s1["occupied_cell_steps"] - s0["occupied_cell_steps"],
            "births": s1["births_total"] - s0["births_total"], "deaths": s1["deaths_total"] - s0["deaths_total"],
            "moves": s1["moves_total"] - s0["moves_total"], "fertile": s1["fertile_total"] - s0["fertile_total"],
            "steps": steps}


TUNING = {}   # launch shape of the plain path (--group-agents, --blocks-per-cu); results do not depend on it


def gpu_run(sc, warmup, steps, timing):
    from qhg4_amd.population import GpuPopulation
    pop = GpuPopulation.from_scenario(sc, capacity=1 << 23)
    if TUNING:
        pop.set_tuning(**TUNING)
    pop.preLoop(0.0)
    if warmup:
        pop.run(warmup)
    pop.sync()
    if timing:
        pop.set_timing(True)
    s0 = pop.stats()
    t0 = time.perf_counter()
    pop.run(steps)
    pop.sync()
    t1 = time.perf_counter()
    s1 = pop.stats()
    times = pop.kernel_times() if timing else None
    pop.close()
    return t1 - t0, s0, s1, times


def host_threads():
    """Threads for the CPU baseline: every core this process may run on
    (its affinity mask), bounded by OMP_NUM_THREADS where the host sets it
    (the share of the host's cores that goes with one GPU)."""
    cores = len(os.sched_getaffinity(0))
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(cores, omp) if omp > 0 else cores), cores


def cpu_baseline(sc, warmup, steps, budget_s):
    from oracle.oracle import Oracle, lib
    threads, cores = host_threads()
    lib().orc_set_threads(threads)
    o = Oracle(sc)
    o.pre_loop()
    if warmup:
        o.run(warmup)
    s0 = o.stats()
    t0 = time.perf_counter()
    done = 0
    while done < steps and time.perf_counter() - t0 < budget_s:
        o.step()
        done += 1
    t1 = time.perf_counter()
    s1 = o.stats()
    agent_steps = s1["agent_steps"] - s0["agent_steps"]
    return {"value": agent_steps / (t1 - t0), "unit": "agent-steps/s", "cores": threads, "kind": "port",
            "sample": "oracle (reference step loop restated in C, OpenMP, same Philox draws) on steps "
                      "%d..%d of the same workload (%d of the %d timed steps, %.1f s, %d agent-steps); "
                      "%d OpenMP threads = the cores in this process's affinity mask (%d) bounded by "
                      "OMP_NUM_THREADS (%s)"
                      % (warmup, warmup + done, done, steps, t1 - t0, agent_steps, threads, cores,
                         os.environ.get("OMP_NUM_THREADS", "unset"))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1900)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c2a", "c2nav", "c3", "c4", "c5"])
    ap.add_argument("--requests", type=int, default=2000, help="C5: requests in the stream")
    ap.add_argument("--budget-gb", type=float, default=250.0, help="C5: HBM budget for resident state")
    ap.add_argument("--chunk", type=int, default=16, help="C5: steps per batch chunk")
    ap.add_argument("--plain-sample", type=int, default=16, help="C5: requests also run alone")
    ap.add_argument("--replicas", type=int, default=256, help="C4: replicas in the ensemble")
    ap.add_argument("--no-ensemble", action="store_true", help="C2 line without the C4 ensemble object")
    ap.add_argument("--no-serving", action="store_true", help="C2 line without the C5 serving object")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="C5 requests for the serving object of the C2 line")
    ap.add_argument("--no-extra-c2", action="store_true",
                    help="C2 line without the full-run and stated-patch objects")
    ap.add_argument("--no-graph", action="store_true", help="C4: launch batch steps without graph replay")
    ap.add_argument("--no-tune", action="store_true", help="C4: keep the default launch shapes")
    ap.add_argument("--group-agents", type=float, default=None, help="plain path: target agents per wave group")
    ap.add_argument("--blocks-per-cu", type=int, default=None, help="plain path: act/scatter blocks per CU")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="also run one plain device copy of BYTES bytes (torch), the reference kernel "
                         "for scaling rocprof FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the population step has no multi-GPU path); "
              "use --gpus 1", file=sys.stderr)
        return 2

    if args.workload == "c5":
        if "--seed" not in sys.argv:
            args.seed = 2000
        print(json.dumps(run_c5(args)))
        return 0
    if args.workload == "c4":
        if "--steps" not in sys.argv:
            args.steps = 1000
        if "--seed" not in sys.argv:
            args.seed = 1000
        print(json.dumps(run_c4(args)))
        return 0

    from qhg4_amd import abi
    if args.group_agents is not None or args.blocks_per_cu is not None:
        TUNING.update(group_agents=args.group_agents or 64.0, blocks_per_cu=args.blocks_per_cu or 16)
    sc, desc = make_workload(args.workload, args.seed)
    print("bench: workload %s, warmup %d, steps %d" % (args.workload, args.warmup, args.steps),
          file=sys.stderr, flush=True)
    wall, s0, s1, _ = gpu_run(sc, args.warmup, args.steps, timing=False)
    d = delta(s1, s0, args.steps)
    value = d["live"] / wall
    print("bench: timed region %.3f s, %d agent-steps" % (wall, d["live"]), file=sys.stderr, flush=True)

    roof = None
    kernels = {}
    if not args.no_kernel_timing:
        _, t0s, t1s, times = gpu_run(sc, args.warmup, args.steps, timing=True)
        if t1s["agent_steps"] != s1["agent_steps"] or t1s["max_id"] != s1["max_id"]:
            raise SystemExit("bench: the timing run diverged from the timed run")
        p = sc["params"]
        ops = {p.action_kind[i] for i in range(p.num_actions)}
        gw = 2 * abi.genome_words(p.gen_genome_size) if abi.ACT_GENETICS in ops else 0
        kb = kernel_bytes(bytes_model(), d, sc["grid"], ops, gw)
        for name, bts in kb.items():
            avg_ms, n = times.get(name, (0.0, 0))
            if not n:
                continue
            per_launch = bts / n
            gbps = per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
            kernels[name] = {"avg_us": round(avg_ms * 1e3, 3), "launches": n,
                             "bytes_per_launch": int(per_launch), "GBps": round(gbps, 1)}
        dom = max(kernels, key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
        kd = kernels[dom]
        ratio = TRAFFIC_PER_ALGO_BYTE.get(dom) if args.workload == "c2" else None
        roof = {"bound": "hbm", "achieved": kd["GBps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(kd["GBps"] / HBM_PEAK_GBPS, 5),
                "traffic": int(ratio * kd["bytes_per_launch"]) if ratio else None, "kernel": dom,
                "bytes_per_launch": kd["bytes_per_launch"], "avg_launch_us": kd["avg_us"]}

    cpu = None if args.no_cpu else cpu_baseline(sc, args.warmup, args.steps, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "cells": int(sc["grid"]["num_cells"]),
                   "initial_agents": int(len(sc["agents"]["id"])), "seed": args.seed,
                   "steps_simulated": args.warmup + args.steps,
                   "mean_live_agents": round(d["live"] / args.steps, 1),
                   "final_live_agents": int(s1["num_alive"])},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    # the ensemble and serving objects ride on the headline (C2) line only
    if args.workload == "c2" and not args.no_extra_c2:
        out["c2_full_run"] = c2_full_run(sc)
        out["c2_stated_patch"] = c2_stated_patch(args.seed)
    if args.workload == "c2" and not args.no_ensemble:
        e = c4_leg(1000, args.replicas, 1000, graph=not args.no_graph, tune=not args.no_tune,
                   kernel_timing=not args.no_kernel_timing)
        e.pop("make")
        out["ensemble"] = e
    if args.workload == "c2" and not args.no_serving:
        e = c5_leg(2000, args.serve_requests, args.budget_gb, args.chunk, args.plain_sample,
                   graph=not args.no_graph)
        e.pop("make")
        e.pop("first")
        out["serving"] = e
    print(json.dumps(out))
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    return 0


def c2_full_run(sc):
    """The whole 2,000-step C2 run from its initial state, timed as one
    region (the headline value times only the requested window)."""
    wall, s0, s1, _ = gpu_run(sc, 0, 2000, timing=False)
    live = s1["agent_steps"] - s0["agent_steps"]
    print("bench: C2 full run %.3f s, %d agent-steps" % (wall, live), file=sys.stderr, flush=True)
    return {"steps": 2000, "value": round(live / wall, 1), "unit": "agent-steps/s",
            "ms_per_step": round(wall * 1e3 / 2000, 5), "mean_live_agents": round(live / 2000, 1),
            "final_live_agents": int(s1["num_alive"])}


# VerhulstVarK values for the stated C2 start: C4's lowest birth and death
# rates (b0 0.05, d0 0.01) and C1's theta (0.01).  At the stated patch (100,000
# agents in 200 cells, N/K about 13) every value set within the configs'
# ranges dies out; these die out last (about step 1,100).
C2_STATED_VERHULST = (0.05, 0.01, 0.01)


def c2_stated_patch(seed):
    """C2 from the initial state as the configuration states it: 100,000
    agents in the 200 land cells around the seed cell."""
    from qhg4_amd import scenarios
    sc = scenarios.c2(seed=seed, patch_cells=200)
    p = sc["params"]
    p.verh_b0, p.verh_d0, p.verh_theta = C2_STATED_VERHULST
    wall, s0, s1, _ = gpu_run(sc, 0, 2000, timing=False)
    live = s1["agent_steps"] - s0["agent_steps"]
    print("bench: C2 stated patch %.3f s, %d agent-steps, %d alive at the end" % (wall, live, s1["num_alive"]),
          file=sys.stderr, flush=True)
    return {"workload": "C2 from 100,000 agents in a 200-cell patch, VerhulstVarK b0 %g d0 %g theta %g"
                        % C2_STATED_VERHULST,
            "steps": 2000, "value": round(live / wall, 1), "unit": "agent-steps/s",
            "ms_per_step": round(wall * 1e3 / 2000, 5), "agent_steps": int(live),
            "final_live_agents": int(s1["num_alive"])}


def copy_calibration(nbytes):
    """One fill (writes nbytes) and one device-to-device copy (reads and
    writes nbytes) through the HIP runtime: kernels of known traffic that
    scale the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    n = C.c_size_t(nbytes)
    for name, rc in (("hipMalloc", hip.hipMalloc(C.byref(src), n)), ("hipMalloc", hip.hipMalloc(C.byref(dst), n)),
                     ("hipMemset", hip.hipMemset(src, 1, n)), ("sync", hip.hipDeviceSynchronize()),
                     ("hipMemcpy", hip.hipMemcpy(dst, src, n, 3)), ("sync", hip.hipDeviceSynchronize())):
        if rc != 0:
            raise SystemExit("copy calibration: %s failed (%d)" % (name, rc))
    hip.hipFree(src)
    hip.hipFree(dst)
    print("bench: copy calibration %d bytes done" % nbytes, file=sys.stderr)


if __name__ == "__main__":
    sys.exit(main())
