"""The device's Fertility, MultiEvaluator and NPPCapacity rules on the cases
pinned by tests/test_oracle_pins.py, and on a grid carrying climate, coast
and water arrays (the Oceania-Miami substitution included): bit-exact with
the oracle."""
import numpy as np
import pytest

from qhg4_amd import abi, scenarios
from qhg4_amd.population import GpuPopulation
from oracle.oracle import Oracle
from tests.helpers import assert_same_state
from tests.test_oracle_pins import (FERTILITY_CASES, NPP_CELLS, ROWS_CUMULATE, ROWS_PLAIN, fertility_scenario,
                                    mul_eval_scenario, npp_scenario)

pytestmark = pytest.mark.gpu


def _bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


def test_fertility_truth_table_on_device():
    print("case: Fertility truth table, 11 agents, 1 step", flush=True)
    sc = fertility_scenario()
    gpu = GpuPopulation.from_scenario(sc)
    gpu.preLoop()
    gpu.run(1)
    want = np.array([c[4] for c in FERTILITY_CASES], dtype=np.uint8)
    assert np.array_equal(gpu.agents()["life"], want)
    gpu.close()


@pytest.mark.parametrize("cumulate0,rows", [(False, ROWS_PLAIN), (True, ROWS_CUMULATE)])
def test_mul_evaluator_on_device(cumulate0, rows):
    print("case: MultiEvaluator MUL on a 3x3 lattice, cumulate=%d" % cumulate0, flush=True)
    sc = mul_eval_scenario(cumulate0)
    gpu = GpuPopulation.from_scenario(sc)
    ora = Oracle(sc)
    gpu.preLoop()
    ora.pre_loop()
    gpu.run(1)
    ora.run(1)
    w = gpu.env_weights()
    assert np.array_equal(_bits(w), _bits(ora.env_weights()))
    for c, want in rows.items():
        np.testing.assert_allclose(w.reshape(9, 5)[c], want, rtol=1e-14, atol=0)
    gpu.close()


def test_npp_branches_on_device():
    print("case: NPPCapacity branches, %d cells" % len(NPP_CELLS), flush=True)
    sc = npp_scenario()
    gpu = GpuPopulation.from_scenario(sc)
    ora = Oracle(sc)
    gpu.preLoop()
    ora.pre_loop()
    k = gpu.capacities()
    assert np.array_equal(_bits(k), _bits(ora.capacities()))
    np.testing.assert_allclose(k, [c[8] for c in NPP_CELLS], rtol=1e-13, atol=0)
    gpu.close()


def climate_scenario(seed=41):
    """The C2 model on a 10,242-cell grid whose environment also carries
    annual temperature and rainfall, coastal flags, water and an Oceania
    region of low NPP (so the Miami substitution applies), with coastal and
    water bonuses switched on."""
    sc = scenarios.c2(seed=seed, level=5, n_agents=20000, patch_cells=800)
    env, grid = sc["env"], sc["grid"]
    rng = np.random.default_rng(seed)
    n = grid["num_cells"]
    lat, lon = env["latitude"], env["longitude"]
    alt = env["altitude"]
    env["temperature"] = 28.0 - 45.0 * np.abs(lat) / 90.0 + rng.normal(0.0, 3.0, n)
    env["rainfall"] = rng.uniform(0.0, 3500.0, n)
    land = alt > 0
    nbr = grid["neighbors"]
    sea_nb = np.zeros(n, dtype=bool)
    for j in range(grid["connectivity"]):
        nb = nbr[:, j]
        sea_nb |= (nb >= 0) & ~land[np.maximum(nb, 0)]
    env["coastal"] = (land & sea_nb).astype(np.uint8)
    env["water"] = np.where(land, rng.uniform(0.0, 0.4, n), 0.0)
    oce = (lon > 115.0) & (lon < 150.0) & (lat > -12.0) & (lat < 1.0)
    npp = env["base_npp"].copy()
    npp[oce & (rng.random(n) < 0.6)] = 0.02
    env["base_npp"] = npp
    p = sc["params"]
    p.npp_min = 0.05
    p.npp_coastal_factor, p.npp_coastal_min_lat, p.npp_coastal_max_lat = 0.1, -40.0, 40.0
    p.npp_water_factor = 0.05
    return sc, oce & land & (npp < p.npp_min)


def test_climate_grid_capacity_weights_and_steps():
    sc, miami = climate_scenario()
    print("case: C2 model on a climate grid, ico level 5, %d Miami cells, 40 steps" % int(miami.sum()), flush=True)
    assert miami.sum() > 0
    gpu = GpuPopulation.from_scenario(sc)
    ora = Oracle(sc)
    gpu.preLoop()
    ora.pre_loop()
    kg, ko = gpu.capacities(), ora.capacities()
    assert np.array_equal(_bits(kg), _bits(ko)), "capacities differ"
    assert (kg[miami] > 0).sum() > miami.sum() // 2   # the Miami NPP gives these cells a capacity
    gpu.run(1)
    ora.run(1)
    assert np.array_equal(_bits(gpu.env_weights()), _bits(ora.env_weights())), "weights differ"
    for k in range(4):
        gpu.run(10 if k else 9)
        ora.run(10 if k else 9)
        assert_same_state(gpu.agents(), ora.agents(), "climate grid step %d" % (10 * (k + 1)))
    gpu.close()
