"""ctypes mirrors of the structs in include/qhg_gpu.h (layout only, no logic).

The C ABI replaces QHG4's per-step population interface (PopBase /
SPopulation, QHG4/core/PopBase.h:16-121); the parameter block carries the
reference's action attributes (e.g. ``Verhulst_b0`` -> ``verh_b0``).
"""
import ctypes as C

import numpy as np

QHG_OK = 0
QHG_ERR_MEMBER = -6
QHG_LIFE_DEAD, QHG_LIFE_ALIVE, QHG_LIFE_FERTILE = 0, 1, 5

ACT_GETOLD = 1
ACT_FERTILITY = 2
ACT_RANDOMPAIR = 3
ACT_VERHULST = 4
ACT_VERHULSTVARK = 5
ACT_RANDOMMOVE = 6
ACT_WEIGHTEDMOVE = 7
ACT_MULTIEVAL = 8
ACT_NPPCAP = 9
ACT_GENETICS = 10
ACT_OLDAGEDEATH = 11
ACT_ATANDEATH = 12
ACT_ASSORTPAIR = 13
ACT_ASSORTPROT = 14
ACT_NAVIGATE = 15
ACT_LAST = ACT_NAVIGATE

# environment events (QHG4/utils_qhg/EventConsts.h) and the drowning rule
EVENT_GEO, EVENT_CLIMATE, EVENT_VEG = 0x01, 0x02, 0x04
EVENT_KILL_WET = 0x100

# reference action names (Action<T>::getActionName), used for same-prio order
ACTION_NAMES = {
    ACT_GETOLD: "GetOld",
    ACT_FERTILITY: "Fertility",
    ACT_RANDOMPAIR: "RandomPair",
    ACT_VERHULST: "Verhulst",
    ACT_VERHULSTVARK: "VerhulstVarK",
    ACT_RANDOMMOVE: "RandomMove",
    ACT_WEIGHTEDMOVE: "WeightedMove",
    ACT_MULTIEVAL: "MultiEvaluator",
    ACT_NPPCAP: "NPPCapacity",
    ACT_GENETICS: "Genetics",
    ACT_OLDAGEDEATH: "OldAgeDeath",
    ACT_ATANDEATH: "ATanDeath",
    ACT_ASSORTPAIR: "AssortativePairHamming",
    ACT_ASSORTPROT: "AssortativePair",
    ACT_NAVIGATE: "Navigate",
}

ME_ADD_SIMPLE, ME_MUL_SIMPLE, ME_MAX_SIMPLE, ME_MIN_SIMPLE = 0, 2, 3, 5
EV_ALTITUDE, EV_CAPACITY, EV_BASE_NPP = 0, 1, 2

MAX_ACTIONS = 16
MAX_EVALS = 4
MAX_PL_POINTS = 64
NAME_LEN = 48

STREAM_PAIR, STREAM_ACT, STREAM_NEWBORN, STREAM_GENETICS, STREAM_ASSORT, STREAM_NAV = 1, 2, 6, 7, 8, 9
ACTW_WMOVE_GO, ACTW_WMOVE_DIR, ACTW_LBIRTH, ACTW_LDEATH, ACTW_RMOVE_GO, ACTW_RMOVE_DIR = 0, 1, 2, 3, 4, 5


class GridDesc(C.Structure):
    _fields_ = [
        ("num_cells", C.c_int32),
        ("connectivity", C.c_int32),
        ("neighbors", C.POINTER(C.c_int32)),
        ("num_neighbors", C.POINTER(C.c_uint8)),
        ("cell_ids", C.POINTER(C.c_int64)),
    ]


class EnvDesc(C.Structure):
    _fields_ = [
        ("altitude", C.POINTER(C.c_double)),
        ("longitude", C.POINTER(C.c_double)),
        ("latitude", C.POINTER(C.c_double)),
        ("water", C.POINTER(C.c_double)),
        ("ice", C.POINTER(C.c_uint8)),
        ("coastal", C.POINTER(C.c_uint8)),
        ("base_npp", C.POINTER(C.c_double)),
        ("temperature", C.POINTER(C.c_double)),
        ("rainfall", C.POINTER(C.c_double)),
    ]


class Params(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("num_actions", C.c_int32),
        ("action_kind", C.c_int32 * MAX_ACTIONS),
        ("action_prio", C.c_int32 * MAX_ACTIONS),
        ("action_name", (C.c_char * NAME_LEN) * MAX_ACTIONS),
        ("fert_min_age", C.c_double),
        ("fert_max_age", C.c_double),
        ("fert_interbirth", C.c_double),
        ("verh_b0", C.c_double),
        ("verh_d0", C.c_double),
        ("verh_theta", C.c_double),
        ("verh_K", C.c_double),
        ("rmove_prob", C.c_double),
        ("wmove_prob", C.c_double),
        ("me_mode", C.c_int32),
        ("me_num", C.c_int32),
        ("me_input", C.c_int32 * MAX_EVALS),
        ("me_cumulate", C.c_int32 * MAX_EVALS),
        ("me_weight", C.c_double * MAX_EVALS),
        ("me_pl_n", C.c_int32 * MAX_EVALS),
        ("me_pl_x", (C.c_double * MAX_PL_POINTS) * MAX_EVALS),
        ("me_pl_v", (C.c_double * MAX_PL_POINTS) * MAX_EVALS),
        ("npp_water_factor", C.c_double),
        ("npp_coastal_factor", C.c_double),
        ("npp_coastal_min_lat", C.c_double),
        ("npp_coastal_max_lat", C.c_double),
        ("npp_min", C.c_double),
        ("npp_max", C.c_double),
        ("npp_k_min", C.c_double),
        ("npp_k_max", C.c_double),
        ("npp_efficiency", C.c_double),
        ("gen_genome_size", C.c_int32),
        ("gen_num_crossover", C.c_int32),
        ("gen_mutation_rate", C.c_double),
        ("oad_max_age", C.c_double),
        ("oad_uncertainty", C.c_double),
        ("atan_max_age", C.c_double),
        ("atan_range", C.c_double),
        ("atan_slope", C.c_double),
        ("assort_cutoff", C.c_double),
        ("assort_permute", C.c_int32),
        ("assort_pad", C.c_int32),
        ("nav_decay", C.c_double),
        ("nav_dist0", C.c_double),
        ("nav_prob0", C.c_double),
        ("nav_min_dens", C.c_double),
        ("nav_bridge_prob", C.c_double),
    ]


class NavDesc(C.Structure):
    """qhg_nav_desc: ports with their destinations and distances, bridges."""
    _fields_ = [
        ("num_ports", C.c_int32),
        ("ports", C.POINTER(C.c_int32)),
        ("dest_offsets", C.POINTER(C.c_int32)),
        ("dest_cells", C.POINTER(C.c_int32)),
        ("dest_dists", C.POINTER(C.c_double)),
        ("sample_dist", C.c_double),
        ("num_bridges", C.c_int32),
        ("bridges", C.POINTER(C.c_int32)),
    ]


def nav_struct(nav):
    """A NavDesc pointing into a navigation dict (ports, dest_offsets,
    dest_cells, dest_dists: numpy arrays; bridges: (B, 2) int32; sample_dist).
    The dict's arrays must outlive the struct."""
    d = NavDesc()
    d.num_ports = len(nav["ports"])
    d.ports = ptr(nav["ports"], C.c_int32)
    d.dest_offsets = ptr(nav["dest_offsets"], C.c_int32)
    d.dest_cells = ptr(nav["dest_cells"], C.c_int32)
    d.dest_dists = ptr(nav["dest_dists"], C.c_double)
    d.sample_dist = float(nav.get("sample_dist", 0.0))
    d.num_bridges = len(nav["bridges"])
    d.bridges = ptr(nav["bridges"], C.c_int32)
    return d


def nav_arrays(ports, bridges=(), sample_dist=0.0):
    """Navigation dict from {port cell: [(dest cell, distance), ...]} and a
    list of (cell, cell) bridges."""
    keys = list(ports)
    off = np.zeros(len(keys) + 1, dtype=np.int32)
    cells, dists = [], []
    for k, c in enumerate(keys):
        for d, dist in ports[c]:
            cells.append(d)
            dists.append(dist)
        off[k + 1] = len(cells)
    return {"ports": np.array(keys, dtype=np.int32), "dest_offsets": off,
            "dest_cells": np.array(cells, dtype=np.int32), "dest_dists": np.array(dists, dtype=np.float64),
            "sample_dist": float(sample_dist),
            "bridges": np.ascontiguousarray(np.array(bridges, dtype=np.int32).reshape(-1, 2))}


class Agents(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("id", C.POINTER(C.c_int64)),
        ("cell", C.POINTER(C.c_int32)),
        ("life", C.POINTER(C.c_uint8)),
        ("gender", C.POINTER(C.c_uint8)),
        ("birth_time", C.POINTER(C.c_double)),
        ("age", C.POINTER(C.c_double)),
        ("last_birth", C.POINTER(C.c_double)),
        ("mate_id", C.POINTER(C.c_int64)),
        ("num_babies", C.POINTER(C.c_int32)),
        ("genome", C.POINTER(C.c_uint64)),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("num_alive", C.c_int64),
        ("high_water", C.c_int64),
        ("capacity", C.c_int64),
        ("max_id", C.c_int64),
        ("step", C.c_int64),
        ("births_last", C.c_int64),
        ("deaths_last", C.c_int64),
        ("moves_last", C.c_int64),
        ("births_total", C.c_int64),
        ("deaths_total", C.c_int64),
        ("moves_total", C.c_int64),
        ("compactions", C.c_int64),
        ("device_error", C.c_int32),
        ("host_error", C.c_int32),
        ("agent_steps", C.c_int64),
        ("slot_steps", C.c_int64),
        ("fertile_total", C.c_int64),
        ("occupied_cell_steps", C.c_int64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ModuleAttrs(C.Structure):
    """qhg_module_attrs: one <module> of a species block and its <param>s."""
    _fields_ = [
        ("name", C.c_char_p),
        ("num_params", C.c_int32),
        ("keys", C.POINTER(C.c_char_p)),
        ("values", C.POINTER(C.c_char_p)),
    ]


class Tuning(C.Structure):
    """qhg_tuning: launch shapes of the step kernels (results never depend on them)."""
    _fields_ = [
        ("group_agents", C.c_double),
        ("blocks_per_cu", C.c_int32),
        ("pad", C.c_int32),
    ]

    def as_dict(self):
        return {"group_agents": self.group_agents, "blocks_per_cu": self.blocks_per_cu}


def ptr(arr, ctype):
    """Pointer to a contiguous numpy array (or NULL for None)."""
    if arr is None:
        return C.POINTER(ctype)()
    assert arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(C.POINTER(ctype))


def genome_words(genome_size):
    """u64 words per strand: GeneUtils::numNucs2Blocks (QHG4/genes/GeneUtils.h:36)."""
    return (genome_size + 31) // 32


AGENT_FIELDS = {
    "id": (np.int64, C.c_int64),
    "cell": (np.int32, C.c_int32),
    "life": (np.uint8, C.c_uint8),
    "gender": (np.uint8, C.c_uint8),
    "birth_time": (np.float64, C.c_double),
    "age": (np.float64, C.c_double),
    "last_birth": (np.float64, C.c_double),
    "mate_id": (np.int64, C.c_int64),
    "num_babies": (np.int32, C.c_int32),
}


def agents_struct(arrays, n):
    """Build an Agents struct pointing into a dict of numpy arrays."""
    a = Agents()
    a.n = n
    for name, (_, ct) in AGENT_FIELDS.items():
        setattr(a, name, ptr(arrays.get(name), ct))
    a.genome = ptr(arrays.get("genome"), C.c_uint64)
    return a


def alloc_agents(n, gwords_total=0):
    """Empty arrays for a download of n agents (genome: n*gwords_total u64)."""
    out = {name: np.zeros(n, dtype=dt) for name, (dt, _) in AGENT_FIELDS.items()}
    if gwords_total:
        out["genome"] = np.zeros(n * gwords_total, dtype=np.uint64)
    return out
