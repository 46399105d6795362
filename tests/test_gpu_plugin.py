"""The PopBase plugin (qhg4_amd/plugin, GpuOoAPopWrapper.so) loaded and driven
by the reference's own executable side: oracle/_ref/plugin_host, compiled
from QHG4's DynPopFactory, ParamProvider2, PopLooper, PopWriter, PopReader
and SCellGrid/Geography/Vegetation (make -C oracle plugin_host).

The host scans the plugin directory, dlopens the Wrapper whose getInfo()
names OoAPop and calls its createPop (QHG4/populations/DynPopFactory.cpp:80-190),
passes the species XML through readSpeciesData, adds agents by addAgent,
steps through PopLooper (QHG4/core/PopLooper.cpp:96-202), raises a geography
event (updateEvent + flushEvents with the grid's Geography rewritten in
place) and writes the population with PopWriter (writeSpeciesDataQDF,
writeAgentDataQDF).  The second case also dumps mid-run and restores the
dump into a new plugin population through PopReader, which then goes on.
The written agents, the per-agent accessors (getAgentID, getAgentCellIndex,
...) and the per-cell counts must equal the oracle's, bit for bit."""
import os
import re
import subprocess

import numpy as np
import pytest

from qhg4_amd import abi, config, qdf, scenarios
from oracle.oracle import Oracle
from tests.helpers import assert_same_state
from tests.test_gpu_popbase import _write_inputs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "oracle", "_ref", "plugin_host")
PLUGINS = os.path.join(ROOT, "qhg4_amd", "plugin")
XML = os.path.join(ROOT, "tests", "golden", "species_ooa.xml")


def _snapshot(path):
    agents, cells = [], {}
    with open(path) as f:
        head = f.readline().split()
        for ln in f:
            p = ln.split()
            if p[0] == "cell":
                cells[int(p[1])] = int(p[2])
            else:
                agents.append((int(p[0]), int(p[1]), int(p[2]), int(p[3]), int(p[4]), float(p[5])))
    return head, agents, cells


@pytest.mark.parametrize("dump_step", [None, 17])
def test_plugin_through_the_reference_host(tmp_path, dump_step):
    print("case: GpuOoAPopWrapper.so in the reference host, ico level 5, 8k agents, 24 steps, geo event "
          "before step 12, dump/restore at %s" % dump_step, flush=True)
    assert os.path.exists(HOST), "build first (make -C oracle plugin_host)"
    assert os.path.exists(os.path.join(PLUGINS, "GpuOoAPopWrapper.so")), "build first (make -C qhg4_amd/plugin)"
    seed = 91
    sc = scenarios.c2(seed=seed, level=5, n_agents=8000, patch_cells=400)
    spec = config.read_species_xml(XML)
    sc["params"] = config.params_from_species(spec, seed)
    steps, geo_step = 24, 12
    rng = np.random.default_rng(5)
    alt2 = np.array(sc["env"]["altitude"], copy=True)
    occupied = np.unique(sc["agents"]["cell"])
    alt2[rng.choice(occupied, len(occupied) // 6, replace=False)] = -30.0
    ice2 = np.zeros(sc["grid"]["num_cells"], dtype=np.uint8)
    ice2[rng.choice(occupied, len(occupied) // 10, replace=False)] = 1
    d = str(tmp_path)
    _write_inputs(d, sc, spec, seed, steps, geo_step, alt2, ice2)
    with open(os.path.join(d, "plan.txt"), "a") as f:
        f.write("seed %d\n" % seed)
        if dump_step:
            f.write("dump %d\n" % dump_step)
    # the reference's XML reader takes no comments
    xml = os.path.join(d, "species.xml")
    with open(XML) as src, open(xml, "w") as dst:
        dst.write(re.sub(r"<!--.*?-->", "", src.read(), flags=re.S))
    r = subprocess.run([HOST, PLUGINS, spec["class"], xml, d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    log = open(os.path.join(d, "log.txt")).read()
    print(log.splitlines()[0], "...", log.splitlines()[-1], flush=True)
    assert "geo event before step %d" % geo_step in log
    if dump_step:
        assert "dumped and restored after step %d" % dump_step in log

    ora = Oracle(sc)
    ora.pre_loop()
    ora.run(geo_step)
    from qhg4_amd._lib import lib
    flags = lib().qhg_params_event_flags(spec["class"].encode(), abi.EVENT_GEO)
    ora.env_event({"altitude": alt2, "ice": ice2}, flags)
    ora.run(steps - geo_step)
    want = ora.agents()
    assert len(want["id"]) > 1000

    # PopWriter's file: the agents as writeAgentDataQDF wrote them
    got = qdf.read_agents(os.path.join(d, "out.qdf"), spec["species_name"] or "sapiens")
    order = np.argsort(got["id"], kind="stable")
    got = {k: v[order] for k, v in got.items()}
    assert_same_state(got, want, "plugin after %d steps" % steps)
    info = qdf.read_info(os.path.join(d, "out.qdf"), spec["species_name"] or "sapiens")
    assert info["max_id"] == ora.stats()["max_id"]

    # the accessors' snapshot (id order) and the per-cell counts
    head, agents, cells = _snapshot(os.path.join(d, "snap.txt"))
    assert int(head[3]) == len(want["id"]) and int(head[4]) == ora.stats()["max_id"]
    a = np.array([x[:5] for x in agents], dtype=np.int64)
    assert np.array_equal(a[:, 0], want["id"])
    assert np.array_equal(a[:, 1], want["cell"]) and np.array_equal(a[:, 2], want["cell"])
    assert np.array_equal(a[:, 3], want["life"]) and np.array_equal(a[:, 4], want["gender"])
    bt = np.array([x[5] for x in agents])
    assert np.array_equal(bt.astype(np.float32), want["birth_time"].astype(np.float32))
    counts = ora.cell_counts()
    assert all(cells[c] == counts[c] for c in range(len(counts)))


def _write_navigation(path, nav):
    with open(path, "w") as f:
        f.write("ports %d\n" % len(nav["ports"]))
        off = nav["dest_offsets"]
        for k, c in enumerate(nav["ports"]):
            ds = ["%d %r" % (nav["dest_cells"][j], float(nav["dest_dists"][j])) for j in range(off[k], off[k + 1])]
            f.write("%d %d %s\n" % (c, off[k + 1] - off[k], " ".join(ds)))
        f.write("bridges %d\n" % len(nav["bridges"]))
        for a, b in nav["bridges"]:
            f.write("%d %d\n" % (a, b))


def test_plugin_navigates_with_the_grids_navigation(tmp_path):
    """GpuOoANavGen2bitPopWrapper.so in the reference host with a Navigation
    on the grid (NavGroupReader's setData / setBridges): the plugin hands the
    sea-ways and bridges to the device at preLoop; a geography event drowns
    agents and changes the current bridges; a navigation event
    (updateEvent(EVENT_ID_NAV) + flushEvents) replaces the sea-ways.  Genomes
    start at zero (the agent lines carry none), in the oracle too."""
    print("case: GpuOoANavGen2bitPopWrapper.so with navigation, ico level 5, 8k agents, 24 steps, geo event "
          "before step 10, nav event before step 16", flush=True)
    assert os.path.exists(HOST), "build first (make -C oracle plugin_host)"
    assert os.path.exists(os.path.join(PLUGINS, "GpuOoANavGen2bitPopWrapper.so")), "build first"
    seed = 95
    sc = scenarios.c2nav(seed=seed, level=5, n_agents=8000, patch_cells=400, prob0=0.15)
    sc["env"].pop("ice")
    sc["agents"]["genome"] = np.zeros_like(sc["agents"]["genome"])
    xml_text = re.sub(r"<!--.*?-->", "", open(XML).read(), flags=re.S)
    xml_text = xml_text.replace('name="OoAPop"', 'name="OoANavGen2bitPop"')
    # one tag per line, as the reference's XML reader takes them
    xml_text = xml_text.replace("<priorities>", '''<module name="Genetics">
    <param name="Genetics_genome_size" value="256"/>
    <param name="Genetics_num_crossover" value="1"/>
    <param name="Genetics_mutation_rate" value="0.001"/>
    <param name="Genetics_initial_muts" value="random"/>
    <param name="Genetics_create_new_genome" value="1"/>
    <param name="Genetics_bits_per_nuc" value="2"/>
  </module>
  <module name="Navigate">
    <param name="Navigate_decay" value="-0.001"/>
    <param name="Navigate_dist0" value="150.0"/>
    <param name="Navigate_prob0" value="0.15"/>
    <param name="Navigate_min_dens" value="0.0"/>
    <param name="Navigate_bridge_prob" value="0.3"/>
  </module>
  <priorities>
    <prio name="Genetics" value="9"/>
    <prio name="Navigate" value="5"/>''')
    xml_text = xml_text.replace('id="NPP">', 'id="NPP">\n      <param name="NPPPref" value="0 0 1.5 1"/>')
    d = str(tmp_path)
    xml = os.path.join(d, "species.xml")
    with open(xml, "w") as f:
        f.write(xml_text)
    spec = config.read_species_xml(xml)
    sc["params"] = config.params_from_species(spec, seed)
    steps, geo_step, nav_step = 24, 10, 16
    rng = np.random.default_rng(9)
    alt2 = np.array(sc["env"]["altitude"], copy=True)
    occupied = np.unique(sc["agents"]["cell"])
    alt2[rng.choice(occupied, len(occupied) // 6, replace=False)] = -30.0
    ice2 = np.zeros(sc["grid"]["num_cells"], dtype=np.uint8)
    ice2[rng.choice(occupied, len(occupied) // 10, replace=False)] = 1
    _write_inputs(d, sc, spec, seed, steps, geo_step, alt2, ice2)
    from tests.test_gpu_navigate import second_navigation
    nav2 = second_navigation(sc, 7)
    _write_navigation(os.path.join(d, "nav.txt"), sc["nav"])
    _write_navigation(os.path.join(d, "nav2.txt"), nav2)
    with open(os.path.join(d, "plan.txt"), "a") as f:
        f.write("seed %d\nnav nav.txt\nnavev %d nav2.txt\n" % (seed, nav_step))
    r = subprocess.run([HOST, PLUGINS, spec["class"], xml, d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    log = open(os.path.join(d, "log.txt")).read()
    assert "geo event before step %d" % geo_step in log and "nav event before step %d" % nav_step in log

    ora = Oracle(sc)
    ora.pre_loop()
    ora.run(geo_step)
    from qhg4_amd._lib import lib
    flags = lib().qhg_params_event_flags(spec["class"].encode(), abi.EVENT_GEO)
    ora.env_event({"altitude": alt2, "ice": ice2}, flags)
    ora.run(nav_step - geo_step)
    ora.set_navigation(nav2)
    ora.run(steps - nav_step)
    want = ora.agents()
    assert len(want["id"]) > 1000 and ora.stats()["moves_total"] > 0
    got = qdf.read_agents(os.path.join(d, "out.qdf"), spec["species_name"] or "sapiens")
    order = np.argsort(got["id"], kind="stable")
    got = {k: v[order] for k, v in got.items()}
    assert_same_state(got, want, "plugin with navigation after %d steps" % steps)
