"""Writes tests/golden/ks_reference_trajectories.json: population and
occupancy trajectories of the unmodified reference (QHG4's tut_SexualPop, run
by oracle/ref_harness, built with `make -C oracle harness` where the
reference tree is present) over 64 seeds, for the statistical comparison with
this build (tests/test_ks_reference.py, tests/test_gpu_ks.py).

Each seed k starts from the initial population of the scenario with seed k,
written by this build's QDF library and read by the reference's PopReader;
the reference draws from its WELL512 streams seeded from k.

    python tests/golden/make_ks_fixtures.py
"""
import json
import multiprocessing as mp
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests.ks_scenarios import SCENARIOS, SEEDS, scenario  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
OUT = os.path.join(HERE, "ks_reference_trajectories.json")


def run_one(job):
    name, seed = job
    from qhg4_amd import qdf
    spec = SCENARIOS[name]
    sc = scenario(name, seed)
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "pop.qdf")
        qdf.write_agents(f, sc["agents"], sc["params"], species="sapiens", pop_class="tut_SexualPop",
                         num_cells=sc["grid"]["num_cells"])
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([HARNESS, f, "sapiens", str(spec["width"]), str(spec["height"]), str(seed),
                            str(spec["steps"])], capture_output=True, text=True, env=env, timeout=1800)
    if r.returncode != 0:
        raise RuntimeError("%s seed %d: ref_harness failed (%d): %s" % (name, seed, r.returncode, r.stderr[-500:]))
    pop, occ = [], []
    for ln in r.stdout.splitlines():
        p = ln.split()
        if len(p) == 3 and all(x.isdigit() for x in p):
            pop.append(int(p[1]))
            occ.append(int(p[2]))
    if len(pop) != spec["steps"]:
        raise RuntimeError("%s seed %d: %d steps read" % (name, seed, len(pop)))
    return name, seed, pop, occ


def main():
    if not os.path.exists(HARNESS):
        raise SystemExit("build the harness first: make -C oracle harness")
    jobs = [(name, s) for name in SCENARIOS for s in SEEDS]
    with mp.Pool(min(8, os.cpu_count() or 1)) as pool:
        res = pool.map(run_one, jobs)
    out = {"generator": "tests/golden/make_ks_fixtures.py (oracle/ref_harness.cpp on QHG4's tut_SexualPop)",
           "seeds": list(SEEDS), "scenarios": {}}
    for name, spec in SCENARIOS.items():
        rows = sorted((r for r in res if r[0] == name), key=lambda r: r[1])
        cps = spec["checkpoints"]
        out["scenarios"][name] = {
            "description": spec["description"], "steps": spec["steps"], "checkpoints": cps,
            "population": [[r[2][c - 1] for c in cps] for r in rows],
            "occupied": [[r[3][c - 1] for c in cps] for r in rows],
            "extinction_step": [next((i + 1 for i, v in enumerate(r[2]) if v == 0), 0) for r in rows],
        }
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
