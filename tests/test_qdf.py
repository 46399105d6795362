"""QDF (HDF5) I/O on the CPU: round trips of agents and parameters through
our files, the reference's member and attribute names in them (h5dump), and
the reference's own QDF files (population and grid) read by name."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from qhg4_amd import abi, config, qdf, scenarios
from qhg4_amd._lib import QhgError, lib
from tests.helpers import assert_same_state

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/QHG4"
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def agents(n, gw=0, seed=1):
    rng = np.random.default_rng(seed)
    a = {"id": rng.permutation(np.arange(5, 5 + 3 * n, 3)).astype(np.int64),
         "cell": rng.integers(0, 500, n).astype(np.int32),
         "life": rng.choice([1, 5], n).astype(np.uint8),
         "gender": rng.integers(0, 2, n).astype(np.uint8),
         "birth_time": rng.uniform(-50, 10, n), "age": rng.uniform(0, 60, n),
         "last_birth": rng.uniform(-5, 10, n), "mate_id": rng.integers(-3, 10**6, n).astype(np.int64),
         "num_babies": rng.integers(0, 7, n).astype(np.int32)}
    if gw:
        a["genome"] = rng.integers(0, 2**63, n * gw, dtype=np.int64).astype(np.uint64)
    return a


def ooa_params(seed=5):
    return config.params_from_xml(os.path.join(HERE, "golden", "species_ooa.xml"), seed=seed)


@pytest.mark.parametrize("n,gw", [(0, 0), (1, 0), (3000, 0), (777, 16)])
def test_agent_round_trip(tmp_path, n, gw):
    p = ooa_params()
    a = agents(n, gw)
    f = str(tmp_path / "pop.qdf")
    qdf.write_agents(f, a, p, pop_class="OoAPop", num_cells=500, step=37, time=37.0, max_id=10**7,
                     genome_words=gw)
    info = qdf.read_info(f)
    assert (info["num_agents"], info["genome_words"], info["num_cells"]) == (n, gw, 500)
    assert (info["step"], info["time"], info["max_id"], info["seed"]) == (37, 37.0, 10**7, 5)
    assert info["class_name"] == "OoAPop" and info["species_name"] == "sapiens" and info["has_params"]
    b = qdf.read_agents(f)
    for k in ("id", "cell", "life", "gender", "birth_time", "age", "last_birth", "mate_id", "num_babies"):
        assert np.array_equal(a[k], b[k]), k
    if gw:
        assert np.array_equal(a["genome"].reshape(n, gw), b["genome"])
    q = qdf.read_params(f)
    assert bytes(q) == bytes(p)


def test_second_species_and_replacement(tmp_path):
    f = str(tmp_path / "two.qdf")
    p = ooa_params()
    qdf.write_agents(f, agents(10), p, species="sapiens", pop_class="OoAPop", num_cells=500)
    qdf.write_agents(f, agents(20, seed=2), p, species="neander", pop_class="OoAPop", num_cells=500, create=False)
    qdf.write_agents(f, agents(30, seed=3), p, species="sapiens", pop_class="OoAPop", num_cells=500, create=False)
    assert qdf.read_info(f, "sapiens")["num_agents"] == 30
    assert qdf.read_info(f, "neander")["num_agents"] == 20
    with pytest.raises(QhgError, match="no population group"):
        qdf.read_info(f, "erectus")


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not available")
def test_reference_names_in_the_file(tmp_path):
    f = str(tmp_path / "names.qdf")
    qdf.write_agents(f, agents(5), ooa_params(), pop_class="OoAPop", num_cells=500)
    out = subprocess.run([H5DUMP, "-H", f], capture_output=True, text=True, cwd=str(tmp_path)).stdout
    for member in ("LifeState", "CellIdx", "CellID", "AgentID", "BirthTime", "Gender", "Age", "LastBirth"):
        assert '"%s"' % member in out, member
    for attr in ("ClassName", "SpeciesName", "NumCells", "PrioInfo", "FuncName", "FuncPrio", "Verhulst_b0",
                 "WeightedMove_prob", "NPPCap_K_max", "OAD_max_age", "Multi_weight_alt", "AltCapPref",
                 "IDGenState"):
        assert attr in out, attr
    # the dump extras under the reference's names (SPopulation.cpp:2184-2195), and no raw
    # parameter struct: the parameters live in the attributes above only
    assert 'DATASET "DeadSpaces"' in out
    assert "QHGGpuParams" not in out and "MaxAgentID" not in out
    assert 'GROUP "Populations"' in out and 'GROUP "sapiens"' in out and 'DATASET "AgentDataSet"' in out
    assert 'ATTRIBUTE "QHG"' in out and 'ATTRIBUTE "Step"' in out and 'ATTRIBUTE "StartTime"' in out


def test_params_from_flat_attribute_map_equal_the_xml():
    """The path a reference-written species group takes (attributes by name
    + PrioInfo) gives the same parameters as the species XML."""
    import ctypes as C
    p = ooa_params(seed=9)
    spec = config.read_species_xml(os.path.join(HERE, "golden", "species_ooa.xml"))
    kv = {}
    for _, prm in spec["modules"]:
        kv.update(prm)
    kv.update({"ClassName": "OoAPop", "NumCells": "500", "Navigate_decay": "0.1"})   # ignored names
    keys = (C.c_char_p * len(kv))(*[k.encode() for k in kv])
    vals = (C.c_char_p * len(kv))(*[v.encode() for v in kv.values()])
    pn = (C.c_char_p * len(spec["priorities"]))(*[n.encode() for n, _ in spec["priorities"]])
    pv = (C.c_int32 * len(spec["priorities"]))(*[v for _, v in spec["priorities"]])
    q = abi.Params()
    assert lib().qhg_params_from_attribute_map(b"OoAPop", len(kv), keys, vals, pn, pv, len(pn), 9, C.byref(q)) == 0
    assert bytes(q) == bytes(p)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference files not present")
def test_reads_the_references_population_file():
    f = os.path.join(REF, "useful_stuff", "pop_location_sap.qdf")
    info = qdf.read_info(f)
    assert info["num_agents"] == 100 and info["num_cells"] == 162 and info["class_name"] == "OoANavRelCompPop"
    a = qdf.read_agents(f)
    # (the file itself repeats one id: 99 distinct ids among its 100 rows)
    assert a["id"].min() >= 0 and info["max_id"] == a["id"].max() == 100
    assert ((a["cell"] >= 0) & (a["cell"] < 162)).all() and set(a["gender"].tolist()) <= {0, 1}
    with pytest.raises(QhgError, match="unknown population class OoANavRelCompPop"):
        qdf.read_params(f)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference files not present")
def test_reads_the_references_grid_file():
    grid, env = qdf.read_grid(os.path.join(REF, "tutorial_data", "grids", "ico32s.qdf"))
    nc = grid["num_cells"]
    assert nc == 10892 and grid["connectivity"] == 6
    assert np.bincount(grid["num_neighbors"]).tolist() == [0, 0, 0, 0, 0, 12, nc - 12]   # 12 pentagons
    nbr = grid["neighbors"]
    for c in range(0, nc, 97):                                 # the neighbour relation is symmetric
        for j in range(grid["num_neighbors"][c]):
            assert c in nbr[nbr[c, j]]
    assert set(env) >= {"longitude", "latitude", "altitude", "water", "ice", "coastal", "base_npp"}
    assert (np.abs(env["latitude"]) <= 90).all() and (np.abs(env["longitude"]) <= 180).all()
    # a model runs on it (oracle): the C2 actions with the file's fields
    sc = scenarios.c2(seed=5, level=4, n_agents=100, patch_cells=10)
    sc["grid"] = dict(grid, xyz=None)
    sc["env"] = env
    cells = np.flatnonzero(env["altitude"] > 0)[:50].astype(np.int32)
    sc["agents"]["cell"] = cells[np.arange(100) % len(cells)]
    from oracle.oracle import Oracle
    o = Oracle(sc)
    o.pre_loop()
    o.run(5)
    assert o.stats()["step"] == 5


@pytest.mark.parametrize("cls,protein", [("OoANavGen2bitAssHamPop", False), ("OoANavGen2bitAssPop", True)])
def test_assortative_params_round_trip(tmp_path, cls, protein):
    """OoANavGen2bitAssHamPop's and OoANavGen2bitAssPop's parameters, the
    pairing's cutoff and permute flag included, come back byte for byte."""
    sc = scenarios.c2(seed=6, level=3, n_agents=200, patch_cells=40, genetics=True, assort=(0.93, 1),
                      protein=protein)
    p = sc["params"]
    f = str(tmp_path / "ass.qdf")
    qdf.write_agents(f, agents(50, 16), p, pop_class=cls, num_cells=500, genome_words=16)
    q = qdf.read_params(f)
    assert (q.assort_cutoff, q.assort_permute) == (0.93, 1)
    assert bytes(q) == bytes(p)
