"""ctypes binding of the CPU oracle (test infrastructure only).

Used by tests/, by __graft_entry__.smoke() as the checker, and by bench.py's
cpu_baseline leg.  Never imported by the qhg4_amd package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from qhg4_amd import abi
from qhg4_amd.scenarios import grid_desc, env_desc

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libqhg_oracle.so")
_lib = None


REFERENCE_TREE = "/root/reference/QHG4"


def build():
    """Compile the oracle; where the reference sources are present, also
    (into oracle/_ref) the reference utility library the pin tests compare
    against, the reference population run by the statistical comparison
    (ref_harness), the reference host that loads the PopBase plugin
    (plugin_host) and the reference's Navigate on small scenes
    (ref_navigate)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir(REFERENCE_TREE):
        subprocess.run(["make", "-s", "-C", HERE, "ref", "harness", "plugin_host", "navigate"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.orc_create.argtypes = [P(abi.GridDesc), P(abi.EnvDesc), P(abi.Params), P(C.c_void_p)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_agents.argtypes = [C.c_void_p, P(abi.Agents)]
        L.orc_num_alive.argtypes = [C.c_void_p]
        L.orc_num_alive.restype = C.c_int64
        L.orc_get_agents.argtypes = [C.c_void_p, P(abi.Agents)]
        L.orc_pre_loop.argtypes = [C.c_void_p]
        L.orc_set_params.argtypes = [C.c_void_p, P(abi.Params)]
        L.orc_env_event.argtypes = [C.c_void_p, P(abi.EnvDesc), C.c_int32]
        L.orc_step.argtypes = [C.c_void_p]
        L.orc_run.argtypes = [C.c_void_p, C.c_int32]
        L.orc_stats.argtypes = [C.c_void_p, P(abi.Stats)]
        L.orc_cell_counts.argtypes = [C.c_void_p, P(C.c_uint32)]
        L.orc_capacities.argtypes = [C.c_void_p, P(C.c_double)]
        L.orc_env_weights.argtypes = [C.c_void_p, P(C.c_double)]
        L.orc_set_threads.argtypes = [C.c_int32]
        L.orc_get_threads.restype = C.c_int32
        L.orc_philox4x32_10.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        L.orc_draw.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32]
        L.orc_draw.restype = C.c_uint32
        L.orc_wrandd.argtypes = [C.c_uint32]
        L.orc_wrandd.restype = C.c_double
        L.orc_wrandr.argtypes = [C.c_double, C.c_double, C.c_uint32]
        L.orc_wrandr.restype = C.c_double
        L.orc_wrandi.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_wrandi.restype = C.c_uint32
        L.orc_multimask.argtypes = [P(C.c_uint32), C.c_int32]
        L.orc_multimask.restype = C.c_uint64
        L.orc_crossover.argtypes = [P(C.c_uint64), P(C.c_uint64), C.c_int32, P(C.c_uint32), C.c_int32]
        L.orc_mutate_nucs.argtypes = [P(C.c_uint64), C.c_int32, C.c_int32, P(C.c_uint32), P(C.c_uint32)]
        L.orc_bino_table.argtypes = [C.c_double, C.c_int32, C.c_double, P(C.c_double), C.c_int32]
        L.orc_bino_table.restype = C.c_int32
        L.orc_bino_getn.argtypes = [P(C.c_double), C.c_int32, C.c_double]
        L.orc_bino_getn.restype = C.c_int32
        L.orc_betai.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_betai.restype = C.c_double
        L.orc_polyline.argtypes = [C.c_int32, P(C.c_double), P(C.c_double), C.c_double]
        L.orc_polyline.restype = C.c_double
        L.orc_md5.argtypes = [P(C.c_uint8), C.c_int64, P(C.c_uint8)]
        L.orc_proteome.argtypes = [P(C.c_uint64), C.c_int32, C.c_int32, P(C.c_uint8), C.c_int32]
        L.orc_proteome.restype = C.c_int32
        L.orc_prot_matches.argtypes = [P(C.c_uint8), C.c_int32, P(C.c_uint8), C.c_int32]
        L.orc_prot_matches.restype = C.c_int32
        L.orc_set_navigation.argtypes = [C.c_void_p, P(abi.NavDesc)]
        L.orc_nav_table.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_double), C.c_int32]
        L.orc_nav_table.restype = C.c_int32
        L.orc_nav_bridges.argtypes = [C.c_void_p, P(C.c_int32), C.c_int32]
        L.orc_nav_bridges.restype = C.c_int32
        L.orc_nav_execute.argtypes = [C.c_void_p, C.c_int32, P(C.c_double), C.c_int32, P(C.c_int32), P(C.c_int32),
                                      C.c_int32]
        L.orc_nav_execute.restype = C.c_int32
        _lib = L
    return _lib


class Oracle:
    """One oracle population built from a scenario dict."""

    def __init__(self, sc):
        L = lib()
        self.sc = sc
        self.grid = sc["grid"]
        self.gdesc = grid_desc(self.grid)
        self.edesc = env_desc(sc["env"])
        self.params = sc["params"]
        h = C.c_void_p()
        rc = L.orc_create(C.byref(self.gdesc), C.byref(self.edesc), C.byref(self.params), C.byref(h))
        if rc != 0:
            raise RuntimeError("orc_create failed %d" % rc)
        self.h = h
        self.gwords = 2 * abi.genome_words(self.params.gen_genome_size) if self._has_genetics() else 0
        a = sc["agents"]
        s = abi.agents_struct(a, len(a["id"]))
        rc = L.orc_set_agents(self.h, C.byref(s))
        if rc != 0:
            raise RuntimeError("orc_set_agents failed %d" % rc)
        if sc.get("nav") is not None:
            self.set_navigation(sc["nav"])

    def _has_genetics(self):
        return any(self.params.action_kind[i] == abi.ACT_GENETICS for i in range(self.params.num_actions))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def pre_loop(self):
        if lib().orc_pre_loop(self.h) != 0:
            raise RuntimeError("orc_pre_loop failed")

    def set_navigation(self, nav):
        """Sea-ways and bridges (see orc_set_navigation)."""
        self.nav = nav
        d = abi.nav_struct(nav)
        if lib().orc_set_navigation(self.h, C.byref(d)) != 0:
            raise RuntimeError("orc_set_navigation refused the navigation")

    def nav_table(self, cell):
        """(targets, cumulative probabilities) of the port at cell, or None."""
        t = np.zeros(4096, dtype=np.int32)
        p = np.zeros(4096, dtype=np.float64)
        n = lib().orc_nav_table(self.h, int(cell), abi.ptr(t, C.c_int32), abi.ptr(p, C.c_double), 4096)
        return None if n < 0 else (t[:n].copy(), p[:n].copy())

    def nav_bridges(self):
        out = np.zeros(2 * 4096, dtype=np.int32)
        n = lib().orc_nav_bridges(self.h, abi.ptr(out, C.c_int32), 4096)
        return [tuple(out[2 * b:2 * b + 2]) for b in range(n)]

    def nav_execute(self, cell, draws):
        """Navigate::execute at cell over the given draws: (registered
        destinations, draws taken)."""
        d = np.ascontiguousarray(np.asarray(draws, dtype=np.float64).reshape(-1))
        dst = np.zeros(4096, dtype=np.int32)
        used = C.c_int32()
        n = lib().orc_nav_execute(self.h, int(cell), abi.ptr(d, C.c_double) if len(d) else None, len(d),
                                  C.byref(used), abi.ptr(dst, C.c_int32), 4096)
        return list(dst[:n]), used.value

    def set_params(self, params):
        if lib().orc_set_params(self.h, C.byref(params)) != 0:
            raise RuntimeError("orc_set_params: the action set differs")
        self.params = params

    def env_event(self, env, events):
        """Environment event between steps (see orc_env_event)."""
        self.env = dict(getattr(self, "env", self.sc["env"]))
        self.env.update(env)
        self.edesc = env_desc(self.env)
        if lib().orc_env_event(self.h, C.byref(self.edesc), int(events)) != 0:
            raise RuntimeError("orc_env_event failed")

    def step(self):
        lib().orc_step(self.h)

    def run(self, n):
        lib().orc_run(self.h, n)

    def stats(self):
        s = abi.Stats()
        lib().orc_stats(self.h, C.byref(s))
        return s.as_dict()

    def agents(self):
        n = lib().orc_num_alive(self.h)
        out = abi.alloc_agents(n, self.gwords)
        s = abi.agents_struct(out, n)
        lib().orc_get_agents(self.h, C.byref(s))
        if self.gwords:
            out["genome"] = out["genome"].reshape(n, self.gwords)
        return out

    def cell_counts(self):
        out = np.zeros(self.grid["num_cells"], dtype=np.uint32)
        lib().orc_cell_counts(self.h, abi.ptr(out, C.c_uint32))
        return out

    def capacities(self):
        out = np.zeros(self.grid["num_cells"], dtype=np.float64)
        lib().orc_capacities(self.h, abi.ptr(out, C.c_double))
        return out

    def env_weights(self):
        out = np.zeros(self.grid["num_cells"] * (self.grid["connectivity"] + 1), dtype=np.float64)
        lib().orc_env_weights(self.h, abi.ptr(out, C.c_double))
        return out
