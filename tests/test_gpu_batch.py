"""Batched stepping on the GPU: every member of a batch ends in exactly the
state it reaches stepped alone (and, through the plain path, the oracle's)."""
import numpy as np
import pytest

from qhg4_amd import abi, scenarios
from qhg4_amd.population import GpuBatch, GpuPopulation
from tests.helpers import assert_same_state, tut_verhulst

pytestmark = pytest.mark.gpu


def mixed_scenarios():
    """Different grids, models, parameters, seeds and sizes."""
    out = []
    out.append(tut_verhulst(scenarios.c1(seed=31, width=48, height=40, n_agents=3000)))
    sc = scenarios.c2(seed=32, level=5, n_agents=15000, patch_cells=600)
    out.append(sc)
    sc = scenarios.c2(seed=33, level=4, n_agents=4000, patch_cells=200, genetics=True)
    out.append(sc)
    sc = scenarios.c2(seed=34, level=4, n_agents=2500, patch_cells=120)
    p = sc["params"]
    p.wmove_prob, p.verh_b0, p.verh_d0 = 0.45, 0.25, 0.05
    out.append(sc)
    sc = scenarios.c1(seed=35, width=16, height=16, n_agents=1)
    out.append(sc)
    sc = tut_verhulst(scenarios.c1(seed=36, width=6, height=6, n_agents=2500))
    rng = np.random.default_rng(36)
    sc["agents"]["cell"] = rng.choice([0, 7, 14, 21], 2500).astype(np.int32)
    sc["params"].verh_K = 2000.0
    out.append(sc)
    return out


def alone(sc, steps):
    pop = GpuPopulation.from_scenario(sc)
    pop.preLoop()
    pop.run(steps)
    a, s = pop.agents(), pop.stats()
    pop.close()
    return a, s


def test_batch_matches_alone():
    print("case: 6 mixed populations, 60 batched steps", flush=True)
    scs = mixed_scenarios()
    ref = [alone(sc, 60) for sc in scs]
    pops = [GpuPopulation.from_scenario(sc) for sc in scs]
    b = GpuBatch()
    for p in pops:
        p.preLoop()
        b.add(p)
    assert len(b) == len(pops)
    b.run(25)
    b.run(35)
    for k, p in enumerate(pops):
        assert_same_state(p.agents(), ref[k][0], "member %d" % k)
        s = p.stats()
        for f in ("num_alive", "max_id", "step", "births_total", "deaths_total", "moves_total", "agent_steps",
                  "fertile_total", "occupied_cell_steps", "births_last", "deaths_last", "moves_last"):
            assert s[f] == ref[k][1][f], (k, f, s[f], ref[k][1][f])
    b.close()
    for p in pops:
        p.close()


def test_batch_join_and_leave():
    """Members join at different batch steps and leave to continue alone."""
    print("case: staggered joins and leaves", flush=True)
    scs = mixed_scenarios()[:4]
    ref = [alone(sc, 50) for sc in scs]
    pops = [GpuPopulation.from_scenario(sc) for sc in scs]
    for p in pops:
        p.preLoop()
    b = GpuBatch()
    b.add(pops[0])
    pops[1].run(7)                 # runs alone first, joins at its step 7
    b.run(7)
    b.add(pops[1])
    b.add(pops[2])                 # joins at its step 0 while others are at 7
    b.run(20)
    b.remove(pops[0])              # leaves at step 27 and continues alone
    pops[0].run(23)
    b.add(pops[3])
    b.run(23)                      # pops[1] at 50, pops[2] at 43, pops[3] at 23
    b.remove(pops[1])
    b.run(7)                       # pops[2] at 50, pops[3] at 30
    b.remove(pops[2])
    b.run(20)                      # pops[3] at 50
    b.close()
    for k, p in enumerate(pops):
        assert p.stats()["step"] == 50, k
        assert_same_state(p.agents(), ref[k][0], "member %d" % k)
        p.close()


def test_batch_capacity_growth_and_stepping_refused():
    """A member outgrowing its buffers inside a batch; step calls on a member
    are refused."""
    print("case: growth inside a batch", flush=True)
    sc = tut_verhulst(scenarios.c1(seed=41, width=64, height=64, n_agents=3000))
    ref, _ = alone(sc, 80)
    pop = GpuPopulation.from_scenario(sc, capacity=1024)
    pop.preLoop()
    b = GpuBatch()
    b.add(pop)
    with pytest.raises(Exception):
        pop.run(1)
    b.run(80)
    assert_same_state(pop.agents(), ref, "grown member")
    b.close()
    pop.close()
