"""Statistical comparison of the oracle with the unmodified reference (CPU).

tests/golden/ks_reference_trajectories.json holds, for 64 seeds, the live
population and occupied-cell count at fixed checkpoints of QHG4's own
tut_SexualPop (oracle/ref_harness.cpp, written by
tests/golden/make_ks_fixtures.py).  The oracle runs the same scenarios from
the same initial agents over the same seeds; its draws are Philox keyed by
agent id rather than WELL512, so each quantity is compared as a distribution:
two-sample Kolmogorov-Smirnov, alpha 0.01 per quantity.  The seeds are fixed,
so the test is deterministic.

A planted mistake shows that the comparison has power: Verhulst's death
rate read as 2 d0 (a plausible slip) fails it.
"""
import json
import os

import numpy as np
import pytest
import scipy.stats

from oracle.oracle import Oracle
from tests.ks_scenarios import ALPHA, SCENARIOS, SEEDS, scenario

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ks_reference_trajectories.json")


def reference():
    with open(FIXTURE) as fh:
        return json.load(fh)


def trajectories(name, run_seed):
    """{quantity: array [seeds, checkpoints]} from run_seed(sc) -> [(pop, occ) per checkpoint]."""
    spec = SCENARIOS[name]
    pop, occ, ext = [], [], []
    for s in SEEDS:
        rows, extinct = run_seed(scenario(name, s), spec)
        pop.append([r[0] for r in rows])
        occ.append([r[1] for r in rows])
        ext.append(extinct)
    return {"population": np.array(pop), "occupied": np.array(occ), "extinction_step": np.array(ext)}


def oracle_run(sc, spec, mutate=None):
    if mutate:
        mutate(sc["params"])
    o = Oracle(sc)
    o.pre_loop()
    rows, step, extinct = [], 0, 0
    cps = set(spec["checkpoints"])
    for step in range(1, spec["steps"] + 1):
        o.step()
        n = o.stats()["num_alive"]
        if n == 0 and extinct == 0:
            extinct = step
        if step in cps:
            cc = o.cell_counts()
            rows.append((int(cc.sum()), int((cc > 0).sum())))
    return rows, extinct


def ks_failures(name, ours, ref):
    """Quantities whose distributions differ at ALPHA: [(quantity, checkpoint, D, p)]."""
    spec = SCENARIOS[name]
    r = ref["scenarios"][name]
    bad = []
    for q in ("population", "occupied"):
        a = ours[q]
        b = np.array(r[q])
        assert a.shape == b.shape == (len(SEEDS), len(spec["checkpoints"]))
        for j, cp in enumerate(spec["checkpoints"]):
            if np.all(a[:, j] == 0) and np.all(b[:, j] == 0):
                continue
            res = scipy.stats.ks_2samp(a[:, j], b[:, j])
            if res.pvalue < ALPHA:
                bad.append((q, cp, float(res.statistic), float(res.pvalue)))
    # survivors count as extinct one step after the last
    ea = np.where(ours["extinction_step"] == 0, spec["steps"] + 1, ours["extinction_step"])
    eb = np.array(r["extinction_step"])
    eb = np.where(eb == 0, spec["steps"] + 1, eb)
    if np.any(ea <= spec["steps"]) or np.any(eb <= spec["steps"]):
        res = scipy.stats.ks_2samp(ea, eb)
        if res.pvalue < ALPHA:
            bad.append(("extinction_step", spec["steps"], float(res.statistic), float(res.pvalue)))
    return bad


def test_fixture_shape():
    ref = reference()
    assert ref["seeds"] == list(SEEDS)
    for name, spec in SCENARIOS.items():
        r = ref["scenarios"][name]
        assert r["checkpoints"] == spec["checkpoints"] and r["steps"] == spec["steps"]
        assert len(r["population"]) == len(SEEDS)
        # occupied cells never exceed the population or the grid
        pop, occ = np.array(r["population"]), np.array(r["occupied"])
        assert np.all(occ <= pop) and np.all(occ <= spec["width"] * spec["height"])


@pytest.mark.parametrize("name", list(SCENARIOS))
def test_oracle_matches_reference_distribution(name):
    ours = trajectories(name, oracle_run)
    bad = ks_failures(name, ours, reference())
    assert not bad, "distributions differ from the reference's: %s" % bad


def test_planted_death_rate_slip_is_detected():
    def slip(p):
        p.verh_d0 = 2.0 * p.verh_d0
    ours = trajectories("c1", lambda sc, spec: oracle_run(sc, spec, slip))
    assert ks_failures("c1", ours, reference())
