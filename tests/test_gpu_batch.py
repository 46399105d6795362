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


def compare_members(pops, ref):
    for k, p in enumerate(pops):
        assert_same_state(p.agents(), ref[k][0], "member %d" % k)
        s = p.stats()
        for f in ("num_alive", "max_id", "step", "births_total", "deaths_total", "moves_total", "agent_steps"):
            assert s[f] == ref[k][1][f], (k, f, s[f], ref[k][1][f])


def test_graph_replay_and_shapes_keep_results():
    """Steps replayed from a captured graph, with the launch shapes changed
    between runs (which recaptures), end where the members end alone."""
    print("case: graph replay, shapes (32,4) (192,32) (64,16), eager tail", flush=True)
    scs = mixed_scenarios()
    ref = [alone(sc, 48) for sc in scs]
    pops = [GpuPopulation.from_scenario(sc) for sc in scs]
    b = GpuBatch()
    for p in pops:
        p.preLoop()
        b.add(p)
    b.set_graph(True)
    b.set_tuning(32, 4)
    b.run(12)
    b.set_tuning(192, 32)
    b.run(12)
    b.remove(pops[4])              # a member leaves: the table changes, recapture
    pops[4].run(12)
    b.set_tuning(64, 16)
    b.run(12)
    b.add(pops[4])
    b.set_graph(False)
    b.run(12)
    b.sync()
    assert b.get_tuning() == {"group_agents": 64.0, "blocks_per_cu": 16}
    compare_members(pops, ref)
    b.close()
    for p in pops:
        p.close()


def test_batch_autotune_keeps_results():
    print("case: batch autotune on 4 members", flush=True)
    scs = mixed_scenarios()[:4]
    pops = [GpuPopulation.from_scenario(sc) for sc in scs]
    b = GpuBatch()
    for p in pops:
        p.preLoop()
        b.add(p)
    b.set_graph(True)
    tuning, taken = b.autotune(2)
    assert taken == 18
    assert 32 <= tuning["group_agents"] <= 192 and tuning["blocks_per_cu"] in (4, 8, 16, 32)
    assert b.get_tuning() == tuning
    b.run(40 - taken)
    ref = [alone(sc, 40) for sc in scs]
    compare_members(pops, ref)
    b.close()
    for p in pops:
        p.close()


def test_population_tuning_and_autotune_keep_results():
    print("case: single population shapes and autotune", flush=True)
    sc = scenarios.c2(seed=37, level=5, n_agents=20000, patch_cells=800)
    ref = alone(sc, 40)
    pop = GpuPopulation.from_scenario(sc)
    pop.preLoop()
    with pytest.raises(Exception):
        pop.set_tuning(0.5, 16)
    with pytest.raises(Exception):
        pop.set_tuning(64, 33)
    pop.set_tuning(16, 2)
    pop.run(5)
    tuning, taken = pop.autotune(1)
    assert taken == 10             # 6 group targets, then all 4 block counts (2 is not one of them)
    assert pop.get_tuning() == tuning
    pop.run(40 - 5 - taken)
    assert_same_state(pop.agents(), ref[0], "autotuned")
    assert pop.stats()["agent_steps"] == ref[1]["agent_steps"]
    pop.close()


def test_device_bytes_matches_allocation():
    """qhg_gpu_device_bytes against the device's free-memory drop."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()

    def free_now():
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
        return free.value

    print("case: device bytes of a 4M-slot C2 population", flush=True)
    sc = scenarios.c2(seed=38, level=6, n_agents=50000, patch_cells=2000)
    f0 = free_now()
    pop = GpuPopulation.from_scenario(sc, capacity=1 << 22)
    f1 = free_now()
    got = pop.device_bytes()
    used = f0 - f1
    assert got > (1 << 22) * 130
    assert abs(used - got) <= 0.05 * got + (64 << 20), (used, got)
    pop.close()
