"""GpuPopulation: Python handle over the C ABI, with the method names of the
reference's population interface (PopBase / SPopulation,
QHG4/core/PopBase.h:16-121, QHG4/core/SPopulation.h:86-150).

Every step runs on the GPU through libqhg4_amd.so; there is no host fallback.
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import lib, check
from .scenarios import grid_desc, env_desc


# reference attribute name -> qhg_params field, for the host copy of the parameters
_FIELD_OF = {
    "Fertility_min_age": "fert_min_age", "Fertility_max_age": "fert_max_age",
    "Fertility_interbirth": "fert_interbirth", "Verhulst_b0": "verh_b0", "Verhulst_d0": "verh_d0",
    "Verhulst_theta": "verh_theta", "Verhulst_K": "verh_K", "RandomMove_prob": "rmove_prob",
    "WeightedMove_prob": "wmove_prob", "OAD_max_age": "oad_max_age", "OAD_uncertainty": "oad_uncertainty",
    "ATanDeath_max_age": "atan_max_age", "ATanDeath_range": "atan_range", "ATanDeath_slope": "atan_slope",
    "Genetics_mutation_rate": "gen_mutation_rate", "AssortativePairHamming_cutoff": "assort_cutoff",
    "AssortativePairHamming_permute": "assort_permute", "AssortativePair_cutoff": "assort_cutoff",
    "AssortativePair_permute": "assort_permute", "Navigate_decay": "nav_decay", "Navigate_dist0": "nav_dist0",
    "Navigate_prob0": "nav_prob0", "Navigate_min_dens": "nav_min_dens", "Navigate_bridge_prob": "nav_bridge_prob",
}


class GpuPopulation:
    def __init__(self, grid, env, params, agents=None, capacity=0):
        self.grid = grid
        self.env = dict(env or {})
        self.params = params
        self._gdesc = grid_desc(grid)
        self._edesc = env_desc(self.env)
        h = C.c_void_p()
        n0 = 0 if agents is None else len(agents["id"])
        cap = int(capacity) if capacity else max(4 * n0, 1 << 20)
        check(lib().qhg_gpu_create(C.byref(self._gdesc), C.byref(self._edesc), C.byref(params), cap,
                                   C.byref(h)), "qhg_gpu_create")
        self.h = h
        self.gwords = 0
        for i in range(params.num_actions):
            if params.action_kind[i] == abi.ACT_GENETICS:
                self.gwords = 2 * abi.genome_words(params.gen_genome_size)
        if agents is not None:
            self.add_agents(agents)

    @classmethod
    def from_scenario(cls, sc, capacity=0):
        pop = cls(sc["grid"], sc["env"], sc["params"], sc["agents"], capacity)
        if sc.get("nav") is not None:
            pop.set_navigation(sc["nav"])
        return pop

    def set_navigation(self, nav):
        """The grid's sea-ways and bridges (Navigation::setData / setBridges):
        before preLoop the grid's data, afterwards EVENT_ID_NAV with its
        flush.  nav: a dict as abi.nav_arrays builds it."""
        self.nav = nav
        d = abi.nav_struct(nav)
        check(lib().qhg_gpu_set_navigation(self.h, C.byref(d)), "qhg_gpu_set_navigation")

    def close(self):
        if getattr(self, "h", None):
            lib().qhg_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # -- agents ----------------------------------------------------------
    def add_agents(self, agents):
        n = len(agents["id"])
        s = abi.agents_struct(agents, n)
        check(lib().qhg_gpu_upload_agents(self.h, C.byref(s)), "qhg_gpu_upload_agents")

    def getNumAgentsEffective(self):
        n = C.c_int64()
        check(lib().qhg_gpu_num_alive(self.h, C.byref(n)), "qhg_gpu_num_alive")
        return n.value

    def agents(self):
        n = self.getNumAgentsEffective()
        out = abi.alloc_agents(n, self.gwords)
        s = abi.agents_struct(out, n)
        check(lib().qhg_gpu_download_agents(self.h, C.byref(s)), "qhg_gpu_download_agents")
        if self.gwords:
            out["genome"] = out["genome"].reshape(n, self.gwords)
        return out

    # -- step phases (PopLooper::doStep order) -----------------------------
    def preLoop(self, t=0.0):
        check(lib().qhg_gpu_pre_loop(self.h, float(t)), "qhg_gpu_pre_loop")

    def initializeStep(self, t):
        check(lib().qhg_gpu_initialize_step(self.h, float(t)), "qhg_gpu_initialize_step")

    def doActions(self, prio, t):
        check(lib().qhg_gpu_do_actions(self.h, int(prio), float(t)), "qhg_gpu_do_actions")

    def finalizeStep(self):
        check(lib().qhg_gpu_finalize_step(self.h), "qhg_gpu_finalize_step")

    def run(self, nsteps):
        check(lib().qhg_gpu_run(self.h, int(nsteps)), "qhg_gpu_run")

    def sync(self):
        check(lib().qhg_gpu_sync(self.h), "qhg_gpu_sync")

    def compactData(self):
        check(lib().qhg_gpu_compact(self.h), "qhg_gpu_compact")

    def modifyAttributes(self, name, value):
        """PopBase::modifyAttributes for an attribute read every step."""
        check(lib().qhg_gpu_modify_attribute(self.h, name.encode(), float(value)), "qhg_gpu_modify_attribute")
        field = _FIELD_OF.get(name)
        if field:
            cur = getattr(self.params, field)
            setattr(self.params, field, int(value) if isinstance(cur, int) else float(value))

    def set_env(self, env):
        self.env.update(env)
        self._edesc = env_desc(self.env)
        check(lib().qhg_gpu_set_env(self.h, C.byref(self._edesc)), "qhg_gpu_set_env")

    def env_event(self, env, events):
        """updateEvent + flushEvents between steps: the given fields replace
        the current ones; with abi.EVENT_KILL_WET the agents of cells under
        water or ice are removed; capacities and weights are recomputed."""
        self.env.update(env)
        self._edesc = env_desc(self.env)
        check(lib().qhg_gpu_env_event(self.h, C.byref(self._edesc), int(events)), "qhg_gpu_env_event")

    # -- observables -------------------------------------------------------
    def stats(self):
        s = abi.Stats()
        check(lib().qhg_gpu_stats(self.h, C.byref(s)), "qhg_gpu_stats")
        return s.as_dict()

    def getNumAgentsArray(self):
        out = np.zeros(self.grid["num_cells"], dtype=np.uint32)
        check(lib().qhg_gpu_cell_counts(self.h, abi.ptr(out, C.c_uint32)), "qhg_gpu_cell_counts")
        return out

    def capacities(self):
        out = np.zeros(self.grid["num_cells"], dtype=np.float64)
        check(lib().qhg_gpu_capacities(self.h, abi.ptr(out, C.c_double)), "qhg_gpu_capacities")
        return out

    def env_weights(self):
        out = np.zeros(self.grid["num_cells"] * (self.grid["connectivity"] + 1), dtype=np.float64)
        check(lib().qhg_gpu_env_weights(self.h, abi.ptr(out, C.c_double)), "qhg_gpu_env_weights")
        return out

    def state_digest(self):
        """Digest of the full state (every live agent's fields and genome),
        independent of the order the device holds the agents in:
        "<live agents>:<hash>:<hash>"."""
        out = (C.c_uint64 * 3)()
        check(lib().qhg_gpu_state_digest(self.h, out), "qhg_gpu_state_digest")
        return "%d:%016x:%016x" % (out[0], out[1], out[2])

    def set_timing(self, on):
        check(lib().qhg_gpu_set_timing(self.h, 1 if on else 0), "qhg_gpu_set_timing")

    def device_bytes(self):
        n = C.c_int64()
        check(lib().qhg_gpu_device_bytes(self.h, C.byref(n)), "qhg_gpu_device_bytes")
        return n.value

    # -- launch shapes (results never depend on them) -----------------------
    def set_tuning(self, group_agents=64.0, blocks_per_cu=16):
        t = abi.Tuning(float(group_agents), int(blocks_per_cu), 0)
        check(lib().qhg_gpu_set_tuning(self.h, C.byref(t)), "qhg_gpu_set_tuning")

    def get_tuning(self):
        t = abi.Tuning()
        check(lib().qhg_gpu_get_tuning(self.h, C.byref(t)), "qhg_gpu_get_tuning")
        return t.as_dict()

    def autotune(self, trial_steps=4):
        """Pick launch shapes on this population's own steps; returns the
        chosen shape and the number of steps the population advanced."""
        t = abi.Tuning()
        n = C.c_int32()
        check(lib().qhg_gpu_autotune(self.h, int(trial_steps), C.byref(t), C.byref(n)), "qhg_gpu_autotune")
        return t.as_dict(), n.value

    def kernel_times(self):
        n = 16
        avg = (C.c_double * n)()
        cnt = (C.c_int64 * n)()
        got = C.c_int32()
        check(lib().qhg_gpu_kernel_times(self.h, avg, cnt, n, C.byref(got)), "qhg_gpu_kernel_times")
        return {lib().qhg_gpu_kernel_name(i).decode(): (avg[i], cnt[i]) for i in range(got.value)}


class GpuBatch:
    """Many populations stepped together: one launch per phase for every
    member (qhg_batch_* in include/qhg_gpu.h).  Each member ends in the state
    it would reach stepped alone; while a member it takes no step calls of
    its own, its observables stay available."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().qhg_batch_create(C.byref(h)), "qhg_batch_create")
        self.h = h
        self.members = []
        self.failed = []      # members that failed a step and left the batch, not yet taken

    def _members_left(self, rc, what):
        """After a run or tuning call: members that failed left the batch
        (their error stays in stats()["host_error"]); the others completed
        the call.  Any other error is the batch's own."""
        if rc == abi.QHG_ERR_MEMBER:
            gone = [p for p in self.members if p.stats()["host_error"] != 0]
            for p in gone:
                self.members.remove(p)
            self.failed.extend(gone)
            return
        check(rc, what)

    def take_failed(self):
        """Members that left the batch on an error since the last call."""
        out, self.failed = self.failed, []
        return out

    def add(self, pop):
        check(lib().qhg_batch_add(self.h, pop.h), "qhg_batch_add")
        self.members.append(pop)

    def remove(self, pop):
        check(lib().qhg_batch_remove(self.h, pop.h), "qhg_batch_remove")
        self.members.remove(pop)

    def __len__(self):
        n = C.c_int32()
        check(lib().qhg_batch_size(self.h, C.byref(n)), "qhg_batch_size")
        return n.value

    def run(self, nsteps):
        """Advance every member by nsteps steps.  A member that fails leaves
        the batch (see take_failed); the others still take every step."""
        self._members_left(lib().qhg_batch_run(self.h, int(nsteps)), "qhg_batch_run")

    def sync(self):
        check(lib().qhg_batch_sync(self.h), "qhg_batch_sync")

    def set_timing(self, on):
        check(lib().qhg_batch_set_timing(self.h, 1 if on else 0), "qhg_batch_set_timing")

    def set_graph(self, on):
        check(lib().qhg_batch_set_graph(self.h, 1 if on else 0), "qhg_batch_set_graph")

    def set_tuning(self, group_agents=64.0, blocks_per_cu=16):
        t = abi.Tuning(float(group_agents), int(blocks_per_cu), 0)
        check(lib().qhg_batch_set_tuning(self.h, C.byref(t)), "qhg_batch_set_tuning")

    def get_tuning(self):
        t = abi.Tuning()
        check(lib().qhg_batch_get_tuning(self.h, C.byref(t)), "qhg_batch_get_tuning")
        return t.as_dict()

    def autotune(self, trial_steps=2):
        """Pick launch shapes on the members' own steps; returns the chosen
        shape and the number of steps every member advanced."""
        t = abi.Tuning()
        n = C.c_int32()
        self._members_left(lib().qhg_batch_autotune(self.h, int(trial_steps), C.byref(t), C.byref(n)),
                           "qhg_batch_autotune")
        return t.as_dict(), n.value

    def kernel_times(self):
        n = 8
        avg = (C.c_double * n)()
        cnt = (C.c_int64 * n)()
        got = C.c_int32()
        check(lib().qhg_batch_kernel_times(self.h, avg, cnt, n, C.byref(got)), "qhg_batch_kernel_times")
        return {lib().qhg_batch_kernel_name(i).decode(): (avg[i], cnt[i]) for i in range(got.value)}

    def close(self):
        if getattr(self, "h", None):
            lib().qhg_batch_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        self.close()
