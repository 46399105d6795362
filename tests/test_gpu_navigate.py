"""Navigate on the GPU against the oracle (bit-exact).

A C2-style coastal scenario (scenarios.c2nav) with sea-way ports, bridges and
ice, with Navigate after WeightedMove (agents that moved do not navigate) and
before it (a later WeightedMove registration overrides a jump).  Every step
the live count and the totals (moves count every registration) equal the
oracle's; every few steps the whole agent state does.  Midway a geography
event raises the water ends of bridges (they become current) and a
navigation event replaces the sea-ways; both populations rebuild their
tables at once.  A dump through the QDF files resumes identically.  The
oracle's Navigate is pinned against the reference's own Navigate in
tests/test_oracle_navigate.py.
"""
import numpy as np
import pytest

from qhg4_amd import abi, qdf, scenarios
from qhg4_amd._lib import QhgError
from qhg4_amd.population import GpuBatch, GpuPopulation
from oracle.oracle import Oracle
from tests.helpers import assert_same_state

pytestmark = pytest.mark.gpu


def same_totals(gpu, ora, label):
    g, o = gpu.stats(), ora.stats()
    for k in ("num_alive", "max_id", "moves_total", "births_total", "deaths_total"):
        assert g[k] == o[k], (label, k, g[k], o[k])
    assert g["device_error"] == 0, label


def second_navigation(sc, seed):
    """Other sea-ways: every port keeps its place, destinations drawn anew;
    the port in the middle (in cell order) gets eight destinations 1 km away,
    whose probabilities add up to more than 1: the table build stops there."""
    rng = np.random.default_rng(seed)
    nav = sc["nav"]
    ports = {}
    cells = np.unique(nav["dest_cells"])
    mid = sorted(nav["ports"])[len(nav["ports"]) // 2]
    for k, c in enumerate(nav["ports"]):
        n = 8 if c == mid else int(rng.integers(0, 5))
        pick = rng.choice(cells, size=min(n, len(cells)), replace=False)
        ports[int(c)] = [(int(d), 1.0 if c == mid else float(rng.uniform(20.0, 800.0))) for d in pick]
    return abi.nav_arrays(ports, [tuple(b) for b in nav["bridges"][::2]], sample_dist=1.0)


@pytest.mark.parametrize("nav_first", [False, True])
def test_navigate_matches_oracle(nav_first):
    print("case: c2nav seed %d nav_first %s" % (61 + nav_first, nav_first), flush=True)
    sc = scenarios.c2nav(seed=61 + nav_first, level=5, n_agents=20000, patch_cells=800, prob0=0.15,
                         nav_first=nav_first)
    gpu = GpuPopulation.from_scenario(sc)
    ora = Oracle(sc)
    gpu.preLoop()
    ora.pre_loop()
    jumps = 0
    for step in range(24):
        before = dict(zip(ora.agents()["id"], ora.agents()["cell"])) if step == 3 else None
        gpu.run(1)
        ora.step()
        same_totals(gpu, ora, "step %d" % step)
        if step % 4 == 3:
            got = gpu.agents()
            assert_same_state(got, ora.agents(), "step %d" % step)
            if before is not None:
                # moves that no neighbour step explains: sea-way jumps and bridges
                nbr = sc["grid"]["neighbors"]
                for i, c in zip(got["id"], got["cell"]):
                    b = before.get(i)
                    if b is not None and c != b and c not in nbr[b]:
                        jumps += 1
        if step == 9:
            alt = sc["env"]["altitude"].copy()
            for a, b in sc["nav"]["bridges"]:
                alt[a] = max(alt[a], 10.0)
                alt[b] = max(alt[b], 10.0)
            gpu.env_event({"altitude": alt}, abi.EVENT_GEO)
            ora.env_event({"altitude": alt}, abi.EVENT_GEO)
            assert len(ora.nav_bridges()) == len(sc["nav"]["bridges"])
        if step == 15:
            nav2 = second_navigation(sc, 5)
            gpu.set_navigation(nav2)
            ora.set_navigation(nav2)
            mid = sorted(sc["nav"]["ports"])[len(sc["nav"]["ports"]) // 2]
            assert ora.nav_table(mid) is None and ora.nav_table(max(sc["nav"]["ports"])) is None
    assert jumps > 0
    assert gpu.stats()["num_alive"] > 1000
    gpu.close()


def test_navigate_grows_capacity_and_resumes_from_a_dump(tmp_path):
    print("case: c2nav seed 64, small capacity, dump at 6", flush=True)
    sc = scenarios.c2nav(seed=64, level=4, n_agents=6000, patch_cells=300, prob0=0.15)
    gpu = GpuPopulation.from_scenario(sc, capacity=6144)
    gpu.preLoop()
    gpu.run(6)
    f = str(tmp_path / "nav.qdf")
    qdf.write_population(gpu, f, pop_class="OoANavGen2bitPop")
    gpu.run(6)
    ref = gpu.agents()
    assert gpu.stats()["capacity"] > 6144
    gpu.close()
    b = GpuPopulation(sc["grid"], sc["env"], qdf.read_params(f))
    b.set_navigation(sc["nav"])
    b.preLoop()
    qdf.restore_population(b, f)
    b.run(6)
    assert_same_state(b.agents(), ref, "resumed")
    ora = Oracle(sc)
    ora.pre_loop()
    ora.run(12)
    assert_same_state(b.agents(), ora.agents(), "resumed vs the oracle")
    p = qdf.read_params(f)
    assert (p.nav_prob0, p.nav_bridge_prob) == (0.15, 0.3)
    assert bytes(p) == bytes(sc["params"])
    b.close()


def test_navigate_needs_its_navigation():
    sc = scenarios.c2nav(seed=65, level=4, n_agents=2000, patch_cells=200)
    nav = sc.pop("nav")
    gpu = GpuPopulation.from_scenario(sc)
    with pytest.raises(QhgError, match="needs the navigation"):
        gpu.preLoop()
    bad = dict(nav)
    bad["ports"] = np.array(list(nav["ports"][:-1]) + [nav["ports"][0]], dtype=np.int32)
    with pytest.raises(QhgError, match="listed twice"):
        gpu.set_navigation(bad)
    gpu.set_navigation(nav)
    gpu.preLoop()
    gpu.run(2)
    assert gpu.stats()["device_error"] == 0
    gpu.close()


def test_navigate_members_of_a_batch_match_alone():
    """Two Navigate populations (Navigate after and before WeightedMove) and
    a plain C2 one stepped as one batch: each ends as it does stepped alone,
    a navigation event on one member midway included."""
    print("case: batch of 2 c2nav members and 1 c2 member, 30 steps, nav event at 12", flush=True)
    scs = [scenarios.c2nav(seed=66, level=4, n_agents=5000, patch_cells=250, prob0=0.15),
           scenarios.c2nav(seed=67, level=5, n_agents=9000, patch_cells=400, prob0=0.15, nav_first=True,
                           genetics=False),
           scenarios.c2(seed=68, level=4, n_agents=4000, patch_cells=200)]
    nav2 = second_navigation(scs[0], 3)

    def alone(sc, k):
        pop = GpuPopulation.from_scenario(sc)
        pop.preLoop()
        pop.run(12)
        if k == 0:
            pop.set_navigation(nav2)
        pop.run(18)
        out = pop.agents(), pop.stats()
        pop.close()
        return out

    ref = [alone(sc, k) for k, sc in enumerate(scs)]
    pops = [GpuPopulation.from_scenario(sc) for sc in scs]
    batch = GpuBatch()
    for p in pops:
        p.preLoop()
        batch.add(p)
    batch.run(12)
    pops[0].set_navigation(nav2)
    batch.run(18)
    for k, p in enumerate(pops):
        assert_same_state(p.agents(), ref[k][0], "member %d" % k)
        s = p.stats()
        for f in ("num_alive", "max_id", "moves_total", "births_total", "deaths_total"):
            assert s[f] == ref[k][1][f], (k, f)
    for p in pops:
        batch.remove(p)
        p.close()


def test_navigate_at_the_bench_size():
    """The bench's c2nav workload (40,962 cells, 100,000 agents on a coast,
    278 ports, 814 destinations, 25 bridges) for 30 steps: the whole state
    equals the oracle's every 10 steps."""
    print("case: c2nav seed 42, level 6, 100k agents, 30 steps", flush=True)
    sc = scenarios.c2nav(seed=42, level=6, n_agents=100000, patch_cells=4000, prob0=0.15, genetics=False)
    gpu = GpuPopulation.from_scenario(sc)
    ora = Oracle(sc)
    gpu.preLoop()
    ora.pre_loop()
    for k in range(3):
        gpu.run(10)
        ora.run(10)
        same_totals(gpu, ora, "step %d" % (10 * k + 10))
        assert_same_state(gpu.agents(), ora.agents(), "step %d" % (10 * k + 10))
    assert ora.stats()["num_alive"] > 50000
    gpu.close()
