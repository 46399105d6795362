"""Dump and resume on the GPU: a population written to a QDF file at step k
and restored into a fresh population continues exactly as the
uninterrupted run (agents, genomes, ids of later newborns)."""
import pytest

from qhg4_amd import qdf, scenarios
from qhg4_amd.population import GpuPopulation
from oracle.oracle import Oracle
from tests.helpers import assert_same_state

pytestmark = pytest.mark.gpu


def run(sc, steps):
    pop = GpuPopulation.from_scenario(sc)
    pop.preLoop()
    pop.run(steps)
    a, s = pop.agents(), pop.stats()
    pop.close()
    return a, s


@pytest.mark.parametrize("case,k,m", [("c2", 15, 15), ("c2", 8, 9), ("c3", 7, 10), ("assort", 9, 8)])
def test_dump_and_resume(tmp_path, case, k, m):
    """case "assort": OoANavGen2bitAssHamPop's AssortativePairHamming pairing,
    whose cutoff and permute flag travel in the file."""
    print("case: %s dump at %d, resume for %d" % (case, k, m), flush=True)
    if case == "c2":
        sc = scenarios.c2(seed=51, level=5, n_agents=20000, patch_cells=800)
    elif case == "assort":
        sc = scenarios.c2(seed=53, level=4, n_agents=6000, patch_cells=300, genetics=True, assort=(0.92, 1))
    else:
        sc = scenarios.c2(seed=52, level=4, n_agents=6000, patch_cells=300, genetics=True)
    ref_a, ref_s = run(sc, k + m)
    f = str(tmp_path / "dump.qdf")
    a = GpuPopulation.from_scenario(sc)
    a.preLoop()
    a.run(k)
    cls = {"c2": "OoAPop", "c3": "OoANavGen2bitPop", "assort": "OoANavGen2bitAssHamPop"}[case]
    qdf.write_population(a, f, pop_class=cls)
    a.close()
    info = qdf.read_info(f)
    assert info["step"] == k and info["time"] == float(k)
    params = qdf.read_params(f)
    assert bytes(params) == bytes(sc["params"])
    b = GpuPopulation(sc["grid"], sc["env"], params)
    b.preLoop()
    qdf.restore_population(b, f)
    b.run(m)
    got = b.agents()
    assert_same_state(got, ref_a, "resumed %s" % case)
    s = b.stats()
    assert (s["num_alive"], s["max_id"], s["step"]) == (ref_s["num_alive"], ref_s["max_id"], ref_s["step"])
    b.close()
    # the resumed run equals the oracle's uninterrupted run
    ora = Oracle(sc)
    ora.pre_loop()
    ora.run(k + m)
    assert_same_state(got, ora.agents(), "resumed %s vs the oracle" % case)
    assert ora.stats()["max_id"] == s["max_id"]


def test_extinction_then_dump_and_resume(tmp_path):
    """A population that dies out (the reference removes it and stops the run,
    QHG4/app/Simulator.cpp:303-314, 378-382): the device path reaches zero
    agents on the oracle's step, further steps are no-ops that still advance
    the clock, and the empty population dumps and restores."""
    print("case: C1 48x48 lattice, 800 agents, until extinct", flush=True)
    sc = scenarios.c1(seed=1, width=48, height=48, n_agents=800)
    gpu = GpuPopulation.from_scenario(sc)
    ora = Oracle(sc)
    gpu.preLoop()
    ora.pre_loop()
    gone = None
    for step in range(1, 301):
        gpu.run(1)
        ora.step()
        n = gpu.getNumAgentsEffective()
        assert n == ora.stats()["num_alive"], step
        if n == 0:
            gone = step
            break
    assert gone is not None and gone == 147
    gpu.run(10)
    s = gpu.stats()
    assert s["num_alive"] == 0 and s["step"] == gone + 10 and s["device_error"] == 0
    assert int(gpu.getNumAgentsArray().sum()) == 0
    f = str(tmp_path / "empty.qdf")
    qdf.write_population(gpu, f, pop_class="tut_SexualPop")
    info = qdf.read_info(f)
    assert info["num_agents"] == 0 and info["step"] == gone + 10
    b = GpuPopulation(sc["grid"], sc["env"], qdf.read_params(f))
    b.preLoop()
    qdf.restore_population(b, f)
    b.run(5)
    assert b.stats()["num_alive"] == 0 and b.stats()["step"] == gone + 15
    gpu.close()
    b.close()


def test_resume_into_the_population_that_wrote_the_dump(tmp_path):
    """A dump restored into the population that wrote it, after that one ran
    on: its clock goes back and the steps repeat exactly.  The status
    snapshots of the first pass are still in the host ring's slots and must
    not stand for the repeated steps (capacity starts small, so growth
    depends on them)."""
    print("case: c2 level 4, dump at 10, run to 23, restore, run to 23 again", flush=True)
    sc = scenarios.c2(seed=54, level=4, n_agents=6000, patch_cells=300)
    pop = GpuPopulation.from_scenario(sc, capacity=6144)
    pop.preLoop()
    pop.run(10)
    f = str(tmp_path / "again.qdf")
    qdf.write_population(pop, f, pop_class="OoAPop")
    pop.run(13)
    ref_a, ref_s = pop.agents(), pop.stats()
    assert ref_s["capacity"] > 6144
    qdf.restore_population(pop, f)
    assert pop.stats()["step"] == 10
    pop.run(13)
    assert_same_state(pop.agents(), ref_a, "repeated steps")
    s = pop.stats()
    for k in ("num_alive", "max_id", "step", "device_error"):
        assert s[k] == ref_s[k], k
    pop.close()
