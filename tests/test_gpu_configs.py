"""Every configuration at the size the bench runs it, on the GPU against the
oracle (the loop of QHG4/core/PopLooper.cpp:166-202).

C2  the bench workload itself (40,962 cells, 100,000 agents, seed 42): 200
    steps in lockstep, every agent field, the per-cell counts, capacities
    and evaluator weights bit-exact.
C3  655,362 cells, 2M agents with 2x256-locus genomes: 10 steps bit-exact,
    then 200 more with size-independent checks (recounted cells, dense new
    ids, ages, and every sampled newborn's strands made of its parents').
C4  the bench ensemble (256 replicas x 1,000 steps): every member's full
    state in the batch equals its state run alone (state digests), and one
    member per grid size equals the oracle.
C5  the first requests of the bench's stream: every served result equals
    the same request run alone (state digests), and the small ones equal the
    oracle.
"""
import numpy as np
import pytest

from qhg4_amd import abi, ensemble, scenarios
from qhg4_amd.population import GpuBatch, GpuPopulation
from qhg4_amd.server import ScenarioServer
from oracle.oracle import Oracle
from tests.helpers import assert_same_state

pytestmark = pytest.mark.gpu

STAT_KEYS = ("num_alive", "max_id", "births_last", "deaths_last", "moves_last", "births_total", "deaths_total",
             "moves_total")


def _bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


def _same(gpu, ora, where):
    assert_same_state(gpu.agents(), ora.agents(), where)
    assert np.array_equal(gpu.getNumAgentsArray(), ora.cell_counts()), where + ": counts"
    gs, os_ = gpu.stats(), ora.stats()
    for k in STAT_KEYS:
        assert gs[k] == os_[k], "%s: stat %s %d vs %d" % (where, k, gs[k], os_[k])
    assert gs["device_error"] == 0


def test_c2_as_benched():
    sc = scenarios.c2()
    print("case: C2 bench workload, 40,962 cells, 100,000 agents, seed 42, 200 steps", flush=True)
    gpu = GpuPopulation.from_scenario(sc)
    ora = Oracle(sc)
    gpu.preLoop()
    ora.pre_loop()
    assert np.array_equal(_bits(gpu.capacities()), _bits(ora.capacities())), "capacities"
    gpu.run(1)
    ora.run(1)
    assert np.array_equal(_bits(gpu.env_weights()), _bits(ora.env_weights())), "evaluator weights"
    _same(gpu, ora, "C2 step 1")
    done = 1
    while done < 200:
        n = 20 if done > 1 else 19
        gpu.run(n)
        ora.run(n)
        done += n
        _same(gpu, ora, "C2 step %d" % done)
    assert gpu.stats()["num_alive"] > 50000
    gpu.close()


def _strand_mismatch(child, a, b):
    """Fewest 2-bit loci of `child` that differ from a one-switch mosaic of
    strands a and b (a before the switch, b after, or the other way round,
    the switch anywhere including the ends).  child, a, b: (n, words)."""
    def loci_differ(x, y):
        d = x ^ y
        d = (d | (d >> np.uint64(1))) & np.uint64(0x5555555555555555)
        # one flag per locus, locus k of word w at bit 2k
        shifts = np.arange(0, 64, 2, dtype=np.uint64)
        return ((d[:, :, None] >> shifts) & np.uint64(1)).reshape(len(x), -1).astype(np.int32)
    da, db = loci_differ(child, a), loci_differ(child, b)
    best = None
    for first, second in ((da, db), (db, da)):
        # mismatches with the switch before locus p: first[:p] + second[p:]
        pre = np.concatenate([np.zeros((len(child), 1), np.int32), np.cumsum(first, axis=1)], axis=1)
        suf = np.concatenate([np.cumsum(second[:, ::-1], axis=1)[:, ::-1], np.zeros((len(child), 1), np.int32)],
                             axis=1)
        m = (pre + suf).min(axis=1)
        best = m if best is None else np.minimum(best, m)
    return best


def _check_births(pre, post, step, gw, limit=4000):
    """Newborns of one step against their parents: mothers are the females
    whose number of babies grew; a newborn is placed in its mother's cell of
    the step (before her own move), and within a cell newborn ids follow the
    mothers' ids (ids are ranked by (cell, mother id)); the father is the
    mother's partner.  Each newborn's first strand must be a one-crossover
    mosaic of the mother's strands and its second of the father's, up to a
    handful of mutated loci (binomial, mean 0.5 per birth).  Returns the
    number of newborns checked.  pre and post are id-ordered states."""
    h = gw // 2
    pid, qid = pre["id"], post["id"]
    newborn = qid > pid.max()
    assert np.all(post["birth_time"][newborn] == float(step))
    j = np.minimum(np.searchsorted(pid, qid), len(pid) - 1)
    known = pid[j] == qid
    mother = known & (post["gender"] == 0) & (post["num_babies"] == pre["num_babies"][j] + 1)
    kids = np.flatnonzero(newborn)
    moms = np.flatnonzero(mother)
    kcell = post["cell"][kids].astype(np.int64)
    mcell = pre["cell"][j[moms]].astype(np.int64)          # the cell she gave birth in
    kids = kids[np.lexsort((qid[kids], kcell))]
    moms = moms[np.lexsort((qid[moms], mcell))]
    kcell, mcell = post["cell"][kids].astype(np.int64), pre["cell"][j[moms]].astype(np.int64)
    # cells where every mother survived the step: as many mothers as newborns
    kc, kn = np.unique(kcell, return_counts=True)
    mc, mn = np.unique(mcell, return_counts=True)
    common, ki, mi = np.intersect1d(kc, mc, return_indices=True)
    good = common[kn[ki] == mn[mi]]
    kids = kids[np.isin(kcell, good)]
    moms = moms[np.isin(mcell, good)]
    assert len(kids) == len(moms)
    dad = np.minimum(np.searchsorted(qid, post["mate_id"][moms]), len(qid) - 1)
    has_dad = qid[dad] == post["mate_id"][moms]
    kids, moms, dad = kids[has_dad][:limit], moms[has_dad][:limit], dad[has_dad][:limit]
    if len(kids):
        g = post["genome"]
        mm = _strand_mismatch(g[kids, :h], g[moms, :h], g[moms, h:])
        mf = _strand_mismatch(g[kids, h:], g[dad, :h], g[dad, h:])
        assert np.all(mm + mf <= 8), "newborn strands not made of their parents' (worst %d loci)" % (mm + mf).max()
    return len(kids)


def test_c3_full_size():
    sc = scenarios.c3()
    print("case: C3 full size, 655,362 cells, 2M agents with genomes: 10 steps vs the oracle", flush=True)
    gpu = GpuPopulation.from_scenario(sc, capacity=8_000_000)
    ora = Oracle(sc)
    gpu.preLoop()
    ora.pre_loop()
    assert np.array_equal(_bits(gpu.capacities()), _bits(ora.capacities())), "capacities"
    for step in range(1, 11):
        gpu.run(1)
        ora.run(1)
        _same(gpu, ora, "C3 step %d" % step)
        print("C3 step %d bit-exact, %d agents" % (step, gpu.stats()["num_alive"]), flush=True)
    del ora
    ncells = sc["grid"]["num_cells"]
    gw = gpu.gwords
    births_checked = 0
    step = 10
    for block in range(8):
        # 25 steps, of which the last is checked against its predecessor
        gpu.run(24)
        step += 24
        pre = gpu.agents()
        max_before = gpu.stats()["max_id"]
        gpu.run(1)
        step += 1
        a = gpu.agents()
        st = gpu.stats()
        counts = gpu.getNumAgentsArray()
        assert st["device_error"] == 0
        assert st["num_alive"] == len(a["id"]) == int(counts.sum())
        assert np.array_equal(counts, np.bincount(a["cell"], minlength=ncells))
        assert np.all(np.diff(a["id"]) > 0)
        newborn = a["id"] > max_before
        assert st["max_id"] == max_before + st["births_last"]
        # new ids are dense: every id of (max_before, max_id] that is missing
        # belongs to a newborn that died in its first step (none can)
        assert np.array_equal(a["id"][newborn], np.arange(max_before + 1, st["max_id"] + 1))
        t = float(step - 1)
        assert np.array_equal(a["age"], t - a["birth_time"])
        assert np.all(a["life"] > 0)
        births_checked += _check_births(pre, a, step - 1, gw)
        print("C3 step %d: %d agents, %d newborn strands checked" % (step, st["num_alive"], births_checked),
              flush=True)
    assert births_checked > 1000
    gpu.close()


def test_c4_bench_ensemble():
    n, steps, seed = 256, 1000, 1000
    print("case: C4 bench ensemble, %d replicas x %d steps: batch vs alone (digests)" % (n, steps), flush=True)
    alone, pops = [], []
    for r in range(n):
        sc = scenarios.c4_replica(r, base_seed=seed, steps=steps)
        p = ensemble.create(sc)
        p.run(steps)
        alone.append(ensemble.summary(p))
        p.close()
        pops.append(ensemble.create(sc))
        if r % 64 == 63:
            print("C4: %d replicas run alone" % (r + 1), flush=True)
    b = GpuBatch()
    for p in pops:
        b.add(p)
    b.set_graph(True)
    _, taken = b.autotune(2)
    b.run(steps - taken)
    b.sync()
    levels = set()
    for r, p in enumerate(pops):
        s = ensemble.summary(p)
        assert s == alone[r], "replica %d: batch %s vs alone %s" % (r, s, alone[r])
        levels.add(scenarios.c4_draw(r, seed)["level"])
    b.close()
    for p in pops:
        p.close()
    assert levels == set(scenarios.C4_LEVELS)
    # one replica per grid size against the oracle (the smallest of each)
    by_level = {}
    for r in range(n):
        d = scenarios.c4_draw(r, seed)
        if d["level"] not in by_level or d["n0"] < scenarios.c4_draw(by_level[d["level"]], seed)["n0"]:
            by_level[d["level"]] = r
    for lv, r in sorted(by_level.items()):
        sc = scenarios.c4_replica(r, base_seed=seed, steps=steps)
        print("C4 replica %d (level %d, %d agents) vs the oracle" % (r, lv, len(sc["agents"]["id"])), flush=True)
        gpu = ensemble.create(sc)
        ora = Oracle(sc)
        ora.pre_loop()
        gpu.run(steps)
        ora.run(steps)
        _same(gpu, ora, "C4 replica %d" % r)
        assert ensemble.summary(gpu) == alone[r]
        gpu.close()


def test_c5_bench_stream_sample():
    seed, nreq = 2000, 40
    reqs = scenarios.c5_requests(2000, base_seed=seed)[:nreq]

    def make(req):
        return scenarios.c5_scenario(req, world_seed=seed)
    print("case: C5, the bench stream's first %d requests served, each vs alone" % nreq, flush=True)
    srv = ScenarioServer(make, budget_bytes=250e9, chunk=16, workers=8)
    res, rep = srv.serve(reqs)
    assert rep["failed"] == 0 and len(res) == nreq
    by = {r["id"]: r for r in res}
    small = []
    for q in reqs:
        sc = make(q)
        p = ensemble.create(sc)
        p.run(q["horizon"])
        s = ensemble.summary(p)
        assert by[q["id"]]["summary"] == s, "request %d differs served and alone" % q["id"]
        if q["n0"] * q["horizon"] <= 40_000_000 and len(small) < 4:
            small.append(q["id"])
            ora = Oracle(sc)
            ora.pre_loop()
            ora.run(q["horizon"])
            _same(p, ora, "C5 request %d alone" % q["id"])
            print("C5 request %d (%d agents, %d steps) equals the oracle" % (q["id"], q["n0"], q["horizon"]),
                  flush=True)
        p.close()
    assert len(small) >= 2
