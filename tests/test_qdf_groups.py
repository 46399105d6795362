"""The QDF functions on groups and datasets the caller opened (the calls a
PopBase makes inside the reference's PopWriter / PopReader), on the CPU:
the caller's HDF5 handles come from libhdf5 through ctypes, in the order
PopWriter::write uses them (QHG4/io/PopWriter.cpp:57-130: species group,
then AgentDataSet created with the population's type, then the additional
data).  Files written this way read back through the path functions and the
other way round."""
import ctypes as C
import os

import numpy as np
import pytest

from qhg4_amd import abi, qdf
from tests.test_qdf import agents, ooa_params

H5F_ACC_RDONLY, H5F_ACC_TRUNC, H5P_DEFAULT = 0, 2, 0


@pytest.fixture(scope="module")
def h5():
    qdf.lib()   # loads libhdf5 through the QDF library's own search path
    path = os.path.join(os.path.dirname(qdf.LIB_PATH), "_hdf5", "libhdf5.so.103")
    if not os.path.exists(path):
        pytest.skip("no HDF5 link (QDF library not built)")
    L = C.CDLL(path)
    i64, cp = C.c_int64, C.c_char_p
    for name, args in (("H5Fcreate", [cp, C.c_uint, i64, i64]), ("H5Fopen", [cp, C.c_uint, i64]),
                       ("H5Gcreate2", [i64, cp, i64, i64, i64]), ("H5Gopen2", [i64, cp, i64]),
                       ("H5Screate_simple", [C.c_int, C.POINTER(C.c_uint64), C.c_void_p]),
                       ("H5Dcreate2", [i64, cp, i64, i64, i64, i64, i64]), ("H5Dopen2", [i64, cp, i64])):
        getattr(L, name).argtypes = args
        getattr(L, name).restype = i64
    for name in ("H5Fclose", "H5Gclose", "H5Sclose", "H5Dclose", "H5Tclose"):
        getattr(L, name).argtypes = [i64]
        getattr(L, name).restype = C.c_int
    return L


def _write_like_popwriter(h5, path, a, p, gw, step, time, max_id):
    Q = qdf.lib()
    f = h5.H5Fcreate(path.encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
    pops = h5.H5Gcreate2(f, b"Populations", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
    g = h5.H5Gcreate2(pops, b"sapiens", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
    assert min(f, pops, g) > 0
    qdf.check(Q.qhg_qdf_write_species_group(g, b"sapiens", b"OoAPop", C.byref(p), 500, step, time, max_id),
              "write_species_group")
    n = len(a["id"])
    t = Q.qhg_qdf_agent_type()
    dims = (C.c_uint64 * 1)(n)
    sp = h5.H5Screate_simple(1, dims, None)
    ds = h5.H5Dcreate2(g, b"AgentDataSet", t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
    assert min(t, sp, ds) > 0
    s = abi.agents_struct(a, n)
    qdf.check(Q.qhg_qdf_write_agent_dataset(ds, C.byref(s), None), "write_agent_dataset")
    if gw:
        qdf.check(Q.qhg_qdf_write_genome(g, abi.ptr(a["genome"], C.c_uint64), n, gw), "write_genome")
    for h, close in ((ds, h5.H5Dclose), (sp, h5.H5Sclose), (t, h5.H5Tclose), (g, h5.H5Gclose),
                     (pops, h5.H5Gclose), (f, h5.H5Fclose)):
        assert close(h) >= 0


@pytest.mark.parametrize("n,gw", [(0, 0), (1, 0), (2500, 0), (600, 16)])
def test_group_writes_read_by_path(h5, tmp_path, n, gw):
    p = ooa_params(seed=11)
    a = agents(n, gw, seed=n + 1)
    f = str(tmp_path / "groups.qdf")
    _write_like_popwriter(h5, f, a, p, gw, step=21, time=21.0, max_id=10**6)
    info = qdf.read_info(f)
    assert (info["num_agents"], info["genome_words"], info["num_cells"]) == (n, gw, 500)
    assert (info["step"], info["time"], info["max_id"], info["seed"]) == (21, 21.0, 10**6, 11)
    assert info["class_name"] == "OoAPop"
    b = qdf.read_agents(f)
    for k in ("id", "cell", "life", "gender", "birth_time", "age", "last_birth", "mate_id", "num_babies"):
        assert np.array_equal(a[k], b[k]), k
    if gw:
        assert np.array_equal(a["genome"].reshape(n, gw), b["genome"])
    assert bytes(qdf.read_params(f)) == bytes(p)


@pytest.mark.parametrize("gw", [0, 16])
def test_path_writes_read_by_group(h5, tmp_path, gw):
    Q = qdf.lib()
    p = ooa_params(seed=12)
    n = 900
    a = agents(n, gw, seed=9)
    f = str(tmp_path / "path.qdf")
    qdf.write_agents(f, a, p, pop_class="OoAPop", num_cells=500, step=5, time=5.0, max_id=4242, genome_words=gw)
    fh = h5.H5Fopen(f.encode(), H5F_ACC_RDONLY, H5P_DEFAULT)
    g = h5.H5Gopen2(fh, b"/Populations/sapiens", H5P_DEFAULT)
    ds = h5.H5Dopen2(g, b"AgentDataSet", H5P_DEFAULT)
    assert min(fh, g, ds) > 0
    q = abi.Params()
    info = qdf.QdfInfo()
    qdf.check(Q.qhg_qdf_read_species_group(g, 0, C.byref(q), C.byref(info)), "read_species_group")
    assert bytes(q) == bytes(p)
    assert (info.num_agents, info.genome_words, info.step, info.time, info.max_id) == (n, gw, 5, 5.0, 4242)
    rows = C.c_int64()
    qdf.check(Q.qhg_qdf_dataset_rows(ds, C.byref(rows)), "dataset_rows")
    assert rows.value == n
    out = abi.alloc_agents(n, gw)
    s = abi.agents_struct(out, n)
    s.genome = None
    qdf.check(Q.qhg_qdf_read_agent_dataset(ds, C.byref(s)), "read_agent_dataset")
    for k in ("id", "cell", "life", "gender", "birth_time", "age", "last_birth", "mate_id", "num_babies"):
        assert np.array_equal(a[k], out[k]), k
    if gw:
        gen = np.zeros(n * gw, dtype=np.uint64)
        qdf.check(Q.qhg_qdf_read_genome(g, abi.ptr(gen, C.c_uint64), n, gw), "read_genome")
        assert np.array_equal(gen, a["genome"])
        assert Q.qhg_qdf_read_genome(g, abi.ptr(gen, C.c_uint64), n, gw + 2) != 0
    # a row count that differs from the dataset's is refused
    short = abi.alloc_agents(n - 1, 0)
    s2 = abi.agents_struct(short, n - 1)
    assert Q.qhg_qdf_read_agent_dataset(ds, C.byref(s2)) != 0
    for h, close in ((ds, h5.H5Dclose), (g, h5.H5Gclose), (fh, h5.H5Fclose)):
        assert close(h) >= 0
