"""The C++ host adapter (include/qhg_population.hpp) as the reference's
driver uses a population: tests/cpp/popbase_driver configures OoAPop from
the species XML's modules (readSpeciesData), adds agents from agent data
lines (addAgent), steps in PopLooper::doStep order (initializeStep, one
doActions per priority, finalizeStep) and raises a geography event between
steps (updateEvent + flushEvents).  Its final agents must equal the
oracle's, bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from qhg4_amd import abi, config, scenarios
from oracle.oracle import Oracle
from tests.helpers import assert_same_state

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "cpp", "_build", "popbase_driver")
XML = os.path.join(HERE, "golden", "species_ooa.xml")


def _write_inputs(d, sc, spec, seed, steps, geo_step, alt2, ice2):
    g = sc["grid"]
    with open(os.path.join(d, "grid.bin"), "wb") as f:
        np.array([g["num_cells"], g["connectivity"]], dtype=np.int32).tofile(f)
        np.ascontiguousarray(g["neighbors"], dtype=np.int32).tofile(f)
        np.ascontiguousarray(g["num_neighbors"], dtype=np.uint8).tofile(f)
    for k in ("altitude", "longitude", "latitude", "base_npp"):
        np.ascontiguousarray(sc["env"][k], dtype=np.float64).tofile(os.path.join(d, k + ".f64"))
    alt2.astype(np.float64).tofile(os.path.join(d, "alt2.f64"))
    ice2.astype(np.uint8).tofile(os.path.join(d, "ice2.u8"))
    with open(os.path.join(d, "species.txt"), "w") as f:
        f.write("class\t%s\nspecies\t%s\nseed\t%d\n" % (spec["class"], spec["species_name"] or "sapiens", seed))
        for name, params in spec["modules"]:
            f.write("module\t%s\n" % name)
            for k, v in params.items():
                f.write("param\t%s\t%s\n" % (k, v))
        for name, prio in spec["priorities"]:
            f.write("prio\t%s\t%d\n" % (name, prio))
    a = sc["agents"]
    with open(os.path.join(d, "agents.txt"), "w") as f:
        for i in range(len(a["id"])):
            f.write("%d\t%d %d %r %d %r %r\n" % (a["cell"][i], a["life"][i], a["id"][i], float(a["birth_time"][i]),
                                                 a["gender"][i], float(a["age"][i]), float(a["last_birth"][i])))
    with open(os.path.join(d, "plan.txt"), "w") as f:
        f.write("steps %d\ngeo %d alt2.f64 ice2.u8\n" % (steps, geo_step))


def _read_out(path):
    with open(path, "rb") as f:
        n = int(np.fromfile(f, dtype=np.int64, count=1)[0])
        out = {}
        for k, dt in (("id", np.int64), ("cell", np.int32), ("life", np.uint8), ("gender", np.uint8),
                      ("birth_time", np.float64), ("age", np.float64), ("last_birth", np.float64),
                      ("mate_id", np.int64), ("num_babies", np.int32)):
            out[k] = np.fromfile(f, dtype=dt, count=n)
    return out


def test_adapter_in_poplooper_order(tmp_path):
    print("case: C++ adapter, OoAPop from species XML, ico level 5, 8k agents, 24 steps, "
          "geography event before step 12", flush=True)
    assert os.path.exists(DRIVER), "build first (python -c 'import __graft_entry__ as g; g.build()')"
    seed = 77
    sc = scenarios.c2(seed=seed, level=5, n_agents=8000, patch_cells=400)
    spec = config.read_species_xml(XML)
    sc["params"] = config.params_from_species(spec, seed)
    steps, geo_step = 24, 12
    rng = np.random.default_rng(3)
    alt2 = np.array(sc["env"]["altitude"], copy=True)
    occupied = np.unique(sc["agents"]["cell"])
    alt2[rng.choice(occupied, len(occupied) // 6, replace=False)] = -30.0
    ice2 = np.zeros(sc["grid"]["num_cells"], dtype=np.uint8)
    ice2[rng.choice(occupied, len(occupied) // 10, replace=False)] = 1
    _write_inputs(str(tmp_path), sc, spec, seed, steps, geo_step, alt2, ice2)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    log = open(os.path.join(str(tmp_path), "log.txt")).read()
    print(log.splitlines()[0], "...", log.splitlines()[-1], flush=True)
    got = _read_out(os.path.join(str(tmp_path), "out.bin"))

    ora = Oracle(sc)
    ora.pre_loop()
    ora.run(geo_step)
    from qhg4_amd._lib import lib
    flags = lib().qhg_params_event_flags(spec["class"].encode(), abi.EVENT_GEO)
    assert flags & abi.EVENT_KILL_WET
    ora.env_event({"altitude": alt2, "ice": ice2}, flags)
    ora.run(steps - geo_step)
    want = ora.agents()
    assert len(want["id"]) > 1000
    assert_same_state(got, want, "adapter after %d steps" % steps)
    assert "geo event before step %d" % geo_step in log
