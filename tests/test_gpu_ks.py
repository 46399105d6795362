"""The MI355X path against the unmodified reference, statistically.

The same comparison as tests/test_ks_reference.py, with the trajectories
taken from the HIP library instead of the oracle: 64 seeds of each scenario,
live population and occupied cells at the checkpoints, extinction step;
two-sample Kolmogorov-Smirnov against QHG4's own tut_SexualPop, alpha 0.01
per quantity.
"""
import pytest

from qhg4_amd.population import GpuPopulation
from tests.ks_scenarios import SCENARIOS
from tests.test_ks_reference import ks_failures, reference, trajectories

pytestmark = pytest.mark.gpu


def gpu_run(sc, spec):
    pop = GpuPopulation.from_scenario(sc)
    try:
        pop.preLoop()
        rows, extinct = [], 0
        cps = set(spec["checkpoints"])
        for step in range(1, spec["steps"] + 1):
            pop.run(1)
            if extinct == 0 and pop.getNumAgentsEffective() == 0:
                extinct = step
            if step in cps:
                cc = pop.getNumAgentsArray()
                rows.append((int(cc.sum()), int((cc > 0).sum())))
        return rows, extinct
    finally:
        pop.close()


@pytest.mark.parametrize("name", list(SCENARIOS))
def test_gpu_matches_reference_distribution(name):
    ours = trajectories(name, gpu_run)
    bad = ks_failures(name, ours, reference())
    assert not bad, "distributions differ from the reference's: %s" % bad
