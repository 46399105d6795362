"""QDF (HDF5) files of populations and grids: ctypes binding of
qhg4_amd/libqhg4_qdf.so (include/qhg_qdf.h)."""
import ctypes as C
import os

import numpy as np

from . import abi
from ._lib import QhgError, lib as gpu_lib

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqhg4_qdf.so")
_lib = None

# every symbol declared in include/qhg_qdf.h
EXPORTS = ["qhg_qdf_last_error", "qhg_qdf_write_agents", "qhg_qdf_write_population", "qhg_qdf_read_info",
           "qhg_qdf_read_params", "qhg_qdf_read_agents", "qhg_qdf_restore_population", "qhg_qdf_read_grid_info",
           "qhg_qdf_read_grid", "qhg_qdf_agent_type", "qhg_qdf_write_species_group",
           "qhg_qdf_read_species_group", "qhg_qdf_dataset_rows", "qhg_qdf_write_agent_dataset",
           "qhg_qdf_read_agent_dataset", "qhg_qdf_write_genome", "qhg_qdf_read_genome"]


class QdfInfo(C.Structure):
    _fields_ = [("num_agents", C.c_int64), ("genome_words", C.c_int32), ("num_cells", C.c_int32),
                ("step", C.c_int64), ("time", C.c_double), ("max_id", C.c_int64), ("seed", C.c_uint64),
                ("has_params", C.c_int32), ("pad", C.c_int32), ("class_name", C.c_char * 64),
                ("species_name", C.c_char * 64)]


class QdfGridInfo(C.Structure):
    _fields_ = [("num_cells", C.c_int32), ("max_neighbors", C.c_int32), ("has_geography", C.c_int32),
                ("has_climate", C.c_int32), ("has_vegetation", C.c_int32), ("pad", C.c_int32)]


def lib():
    global _lib
    if _lib is None:
        gpu_lib()                       # the step library first (this one links to it)
        if not os.path.exists(LIB_PATH):
            raise QhgError("%s not built (no HDF5 at build time?)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P, h = C.POINTER, C.c_void_p
        sig = {
            "qhg_qdf_last_error": ([], C.c_char_p),
            "qhg_qdf_write_agents": ([C.c_char_p, C.c_int32, C.c_char_p, C.c_char_p, P(abi.Params), C.c_int32,
                                      P(C.c_int64), C.c_int64, C.c_double, C.c_int64, P(abi.Agents), C.c_int32],
                                     C.c_int),
            "qhg_qdf_write_population": ([h, C.c_char_p, C.c_int32, C.c_char_p, C.c_char_p, C.c_int32,
                                          P(C.c_int64)], C.c_int),
            "qhg_qdf_read_info": ([C.c_char_p, C.c_char_p, P(QdfInfo)], C.c_int),
            "qhg_qdf_read_params": ([C.c_char_p, C.c_char_p, C.c_uint64, P(abi.Params)], C.c_int),
            "qhg_qdf_read_agents": ([C.c_char_p, C.c_char_p, P(abi.Agents)], C.c_int),
            "qhg_qdf_restore_population": ([C.c_char_p, C.c_char_p, h], C.c_int),
            "qhg_qdf_read_grid_info": ([C.c_char_p, P(QdfGridInfo)], C.c_int),
            "qhg_qdf_read_grid": ([C.c_char_p, P(C.c_int32), P(C.c_uint8), P(C.c_int32)] + [P(C.c_double)] * 4 +
                                  [P(C.c_uint8)] * 2 + [P(C.c_double)] * 3 + [P(C.c_int32)], C.c_int),
            # on groups / datasets opened by the caller (HDF5 handles as int64)
            "qhg_qdf_agent_type": ([], C.c_int64),
            "qhg_qdf_write_species_group": ([C.c_int64, C.c_char_p, C.c_char_p, P(abi.Params), C.c_int32, C.c_int64,
                                             C.c_double, C.c_int64], C.c_int),
            "qhg_qdf_read_species_group": ([C.c_int64, C.c_uint64, P(abi.Params), P(QdfInfo)], C.c_int),
            "qhg_qdf_dataset_rows": ([C.c_int64, P(C.c_int64)], C.c_int),
            "qhg_qdf_write_agent_dataset": ([C.c_int64, P(abi.Agents), P(C.c_int64)], C.c_int),
            "qhg_qdf_read_agent_dataset": ([C.c_int64, P(abi.Agents)], C.c_int),
            "qhg_qdf_write_genome": ([C.c_int64, P(C.c_uint64), C.c_int64, C.c_int32], C.c_int),
            "qhg_qdf_read_genome": ([C.c_int64, P(C.c_uint64), C.c_int64, C.c_int32], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        raise QhgError("%s failed (%d): %s" % (what, rc, lib().qhg_qdf_last_error().decode()))


def _b(s):
    return None if s is None else s.encode()


def write_agents(path, agents, params, species="sapiens", pop_class="", num_cells=0, step=0, time=0.0,
                 max_id=None, genome_words=0, create=True, cell_ids=None):
    n = len(agents["id"])
    if max_id is None:
        max_id = int(agents["id"].max()) if n else 0
    s = abi.agents_struct(agents, n)
    cid = None if cell_ids is None else np.ascontiguousarray(cell_ids, dtype=np.int64)
    check(lib().qhg_qdf_write_agents(_b(path), 1 if create else 0, _b(species), _b(pop_class), C.byref(params),
                                     int(num_cells), abi.ptr(cid, C.c_int64), int(step), float(time), int(max_id),
                                     C.byref(s), int(genome_words)), "qhg_qdf_write_agents")


def write_population(pop, path, species="sapiens", pop_class="", create=True, cell_ids=None):
    cid = None if cell_ids is None else np.ascontiguousarray(cell_ids, dtype=np.int64)
    check(lib().qhg_qdf_write_population(pop.h, _b(path), 1 if create else 0, _b(species), _b(pop_class),
                                         int(pop.grid["num_cells"]), abi.ptr(cid, C.c_int64)),
          "qhg_qdf_write_population")


def read_info(path, species="sapiens"):
    i = QdfInfo()
    check(lib().qhg_qdf_read_info(_b(path), _b(species), C.byref(i)), "qhg_qdf_read_info")
    d = {f: getattr(i, f) for f, _ in QdfInfo._fields_ if f != "pad"}
    d["class_name"] = i.class_name.decode()
    d["species_name"] = i.species_name.decode()
    return d


def read_params(path, species="sapiens", seed=0):
    p = abi.Params()
    check(lib().qhg_qdf_read_params(_b(path), _b(species), int(seed), C.byref(p)), "qhg_qdf_read_params")
    return p


def read_agents(path, species="sapiens"):
    info = read_info(path, species)
    n, gw = info["num_agents"], info["genome_words"]
    out = abi.alloc_agents(n, gw)
    s = abi.agents_struct(out, n)
    check(lib().qhg_qdf_read_agents(_b(path), _b(species), C.byref(s)), "qhg_qdf_read_agents")
    if gw:
        out["genome"] = out["genome"].reshape(n, gw)
    return out


def restore_population(pop, path, species="sapiens"):
    check(lib().qhg_qdf_restore_population(_b(path), _b(species), pop.h), "qhg_qdf_restore_population")


def read_grid(path):
    """(grid dict, env dict) of a QDF grid file; the env holds the fields the
    file has."""
    gi = QdfGridInfo()
    check(lib().qhg_qdf_read_grid_info(_b(path), C.byref(gi)), "qhg_qdf_read_grid_info")
    nc, mx = gi.num_cells, gi.max_neighbors
    ids = np.zeros(nc, np.int32)
    nn = np.zeros(nc, np.uint8)
    nbr = np.zeros((nc, mx), np.int32)
    names = ["longitude", "latitude", "altitude", "water", "ice", "coastal", "base_npp", "temperature", "rainfall"]
    arrs = {k: np.zeros(nc, np.uint8 if k in ("ice", "coastal") else np.float64) for k in names}
    miss = C.c_int32()
    ptrs = [abi.ptr(arrs[k], C.c_uint8 if k in ("ice", "coastal") else C.c_double) for k in names]
    check(lib().qhg_qdf_read_grid(_b(path), abi.ptr(ids, C.c_int32), abi.ptr(nn, C.c_uint8),
                                  abi.ptr(nbr, C.c_int32), *ptrs, C.byref(miss)), "qhg_qdf_read_grid")
    env = {k: arrs[k] for i, k in enumerate(names) if not (miss.value >> i) & 1}
    grid = {"num_cells": nc, "connectivity": mx, "neighbors": nbr, "num_neighbors": nn,
            "cell_ids": ids.astype(np.int64)}
    if "longitude" in env:
        grid["lon"], grid["lat"] = env["longitude"], env["latitude"]
    return grid, env
