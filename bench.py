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

# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on the
# default workload and steps (100 warm-up + 1,900 timed), each counter scaled by
# what it reads for a 256 MiB device copy (FETCH x2, WRITE x1); see
# profiles/step_kernels_c2_counters.md.  Reported as roofline.traffic only for
# that exact workload.
MEASURED_TRAFFIC = {("c2", 100, 1900): {"k_cell_act": 22.28e6, "k_cell_scatter": 41.93e6, "k_cell_scan": 7.14e6}}
METRIC = "agent-steps/sec at 1 MI355X + %HBM peak; ensemble vs plain speedup, p50/p99"


def bytes_model():
    path = os.path.join(ROOT, "qhg4_amd", "csrc", "bytes_model.h")
    out = {}
    for ln in open(path):
        m = re.match(r"#define\s+QHG_B_(\w+)\s+(\d+)", ln)
        if m:
            out[m.group(1)] = int(m.group(2))
    return out


def grid_edges(grid):
    """Incoming (source cell, neighbour slot) pairs a move can take."""
    import numpy as np
    nbr = np.asarray(grid["neighbors"]).reshape(grid["num_cells"], grid["connectivity"])
    nn = np.asarray(grid["num_neighbors"])
    slot = np.arange(grid["connectivity"])[None, :]
    return int(((nbr >= 0) & (slot < nn[:, None])).sum())


def kernel_bytes(bm, d, grid, ops, gwords):
    """Algorithmic bytes of each step kernel summed over the measured steps.
    d: deltas of live agent-steps, occupied cell-steps, births, deaths."""
    from qhg4_amd import abi
    ncells, conn = grid["num_cells"], grid["connectivity"]
    live, occ, births, deaths, steps = d["live"], d["occupied"], d["births"], d["deaths"], d["steps"]
    per_live = bm["ACT_LIVE_BASE"]
    if abi.ACT_RANDOMPAIR in ops:
        per_live += bm["ACT_LIVE_PAIR"]
    if abi.ACT_GETOLD in ops:
        per_live += bm["ACT_LIVE_GETOLD"]
    if abi.ACT_FERTILITY in ops:
        per_live += bm["ACT_LIVE_FERT"]
    per_cell = bm["ACT_CELL"] + conn * bm["ACT_CELL_PER_CONN"]
    per_occ = bm["ACT_OCC"] + conn * bm["ACT_OCC_PER_CONN"] + bm["ACT_OCC_W_EXTRA"]
    act = steps * ncells * per_cell + occ * per_occ + live * per_live + births * bm["ACT_BIRTH"]
    scan = steps * (ncells * bm["SCAN_CELL"] + grid_edges(grid) * bm["SCAN_EDGE"])
    survivors = live - deaths
    per_surv = bm["SCAT_SURVIVOR"] + (bm["SCAT_SURVIVOR_GEN"] if gwords else 0)
    per_birth = bm["SCAT_BIRTH"] + (bm["SCAT_BIRTH_GEN"] if gwords else 0)
    scat = steps * ncells * bm["SCAT_CELL"] + live * bm["SCAT_LIVE"] + survivors * per_surv + births * per_birth
    out = {"k_cell_act": act, "k_cell_scan": scan, "k_cell_scatter": scat}
    if gwords:
        out["k_genome"] = births * bm["SCAT_GENOME_WORDS"] * gwords * 8
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
    elif name == "c3":
        sc = scenarios.c3(seed=seed)
        desc = "C3: C2 model + 2x256-locus 2-bit genetics on a 655,362-cell icosahedral grid, 2M agents"
    else:
        raise SystemExit("unknown workload %s" % name)
    return sc, desc


def delta(s1, s0, steps):
    return {"live": s1["agent_steps"] - s0["agent_steps"],
            "occupied": s1["occupied_cell_steps"] - s0["occupied_cell_steps"],
            "births": s1["births_total"] - s0["births_total"], "deaths": s1["deaths_total"] - s0["deaths_total"],
            "moves": s1["moves_total"] - s0["moves_total"], "steps": steps}


def gpu_run(sc, warmup, steps, timing):
    from qhg4_amd.population import GpuPopulation
    pop = GpuPopulation.from_scenario(sc, capacity=1 << 23)
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


def cpu_baseline(sc, warmup, steps, budget_s):
    from oracle.oracle import Oracle, lib
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
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
                      "%d..%d of the same workload (%d of the %d timed steps, %.1f s, %d agent-steps)"
                      % (warmup, warmup + done, done, steps, t1 - t0, agent_steps)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1900)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3"])
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

    from qhg4_amd import abi
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
        traffic = MEASURED_TRAFFIC.get((args.workload, args.warmup, args.steps), {}).get(dom)
        roof = {"bound": "hbm", "achieved": kd["GBps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(kd["GBps"] / HBM_PEAK_GBPS, 5), "traffic": traffic, "kernel": dom,
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
    print(json.dumps(out))
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    return 0


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
