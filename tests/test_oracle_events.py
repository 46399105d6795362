"""Environment events in the oracle (orc_env_event), pinned against facts that
do not depend on its code: the drowning rule of OoAPop::updateEvent
(QHG4/populations/OoAPop.cpp:164-196) removes exactly the agents of cells
with altitude < 0 or ice and touches no other agent; the recomputed
capacities and weights equal those of a population built on the new
environment from the start."""
import copy

import numpy as np
import pytest

from qhg4_amd import abi, scenarios
from qhg4_amd._lib import lib
from oracle.oracle import Oracle


def _scenario(genetics=False):
    return scenarios.c2(seed=13, level=4, n_agents=4000, patch_cells=200, genetics=genetics)


def _wet_env(sc, cells_down, cells_ice):
    env = {k: np.array(v, copy=True) for k, v in sc["env"].items()}
    env["altitude"][cells_down] = -25.0
    ice = np.zeros(sc["grid"]["num_cells"], dtype=np.uint8)
    ice[cells_ice] = 1
    env["ice"] = ice
    return env


@pytest.mark.parametrize("genetics", [False, True])
def test_kill_wet_removes_exactly_the_wet_cells(genetics):
    sc = _scenario(genetics)
    ora = Oracle(sc)
    ora.pre_loop()
    ora.run(6)
    before = ora.agents()
    occupied = np.unique(before["cell"])
    rng = np.random.default_rng(4)
    pick = rng.permutation(occupied)
    down, ice = pick[:len(pick) // 4], pick[len(pick) // 4: len(pick) // 3]
    env = _wet_env(sc, down, ice)
    max_id = ora.stats()["max_id"]
    ora.env_event(env, abi.EVENT_GEO | abi.EVENT_KILL_WET)
    after = ora.agents()
    wet = np.isin(before["cell"], np.concatenate([down, ice]))
    assert wet.sum() > 0 and (~wet).sum() > 0
    assert len(after["id"]) == int((~wet).sum())
    for k in ("id", "cell", "life", "gender", "birth_time", "age", "last_birth", "num_babies"):
        assert np.array_equal(after[k], before[k][~wet]), k
    assert np.all(after["mate_id"] == -3)
    if genetics:
        assert np.array_equal(after["genome"], before["genome"][~wet])
    counts = ora.cell_counts()
    assert counts[down].sum() == 0 and counts[ice].sum() == 0
    assert counts.sum() == len(after["id"])
    s = ora.stats()
    assert s["max_id"] == max_id and s["num_alive"] == len(after["id"])
    # stepping goes on from the survivors
    ora.run(3)
    assert ora.stats()["num_alive"] > 0


def test_geo_event_without_kill_keeps_agents():
    sc = _scenario()
    ora = Oracle(sc)
    ora.pre_loop()
    ora.run(4)
    before = ora.agents()
    occupied = np.unique(before["cell"])
    ora.env_event(_wet_env(sc, occupied[:10], occupied[10:20]), abi.EVENT_GEO)
    after = ora.agents()
    assert np.array_equal(after["id"], before["id"])


def test_recomputed_tables_match_a_fresh_build():
    sc = _scenario()
    ora = Oracle(sc)
    ora.pre_loop()
    ora.run(3)
    k_before = ora.capacities()
    env = {k: np.array(v, copy=True) for k, v in sc["env"].items()}
    env["base_npp"] = env["base_npp"] * 0.6
    env["altitude"] = env["altitude"] + 150.0
    ora.env_event(env, abi.EVENT_CLIMATE | abi.EVENT_GEO)
    fresh_sc = copy.copy(sc)
    fresh_sc["env"] = {k: np.array(v, copy=True) for k, v in env.items()}
    fresh = Oracle(fresh_sc)
    fresh.pre_loop()
    fresh.run(1)   # the evaluator weights are built in the first step's initialize
    assert np.array_equal(ora.capacities().view(np.uint64), fresh.capacities().view(np.uint64))
    assert np.array_equal(ora.env_weights().view(np.uint64), fresh.env_weights().view(np.uint64))
    assert not np.array_equal(ora.capacities(), k_before)


def test_event_flags_by_class():
    L = lib()
    geo = abi.EVENT_GEO
    for cls in (b"tut_EnvironCapAltPop", b"OoAPop", b"OoANavGen2bitPop"):
        assert L.qhg_params_event_flags(cls, geo) == geo | abi.EVENT_KILL_WET
        assert L.qhg_params_event_flags(cls, abi.EVENT_VEG) == abi.EVENT_VEG
    assert L.qhg_params_event_flags(b"tut_SexualPop", geo) == geo
    assert L.qhg_params_event_flags(b"NoSuchPop", geo) < 0
