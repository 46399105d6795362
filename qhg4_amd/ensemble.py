"""Ensembles of independent scenario runs on one GPU.

run_plain    every scenario alone through the plain step path, one after
             another (create, pre-loop, run, read back, destroy).
run_batched  every scenario as a member of one batch: one launch per phase
             steps all members; optionally each step is replayed from a
             captured HIP graph and the launch shapes are picked by the
             autotuner on the first steps.

Both give every scenario exactly the state it reaches alone (the batch kernels
use the members' own Philox keys and step counters), so their per-scenario
summaries, which carry a digest of the full agent state, are compared for
equality.  There is no counterpart in the
reference, which runs one population per process (QHG4/core/PopLooper.cpp).
"""
import time

from .population import GpuBatch, GpuPopulation

SUMMARY_KEYS = ("num_alive", "max_id", "step", "births_total", "deaths_total", "moves_total", "agent_steps",
                "fertile_total", "occupied_cell_steps")


def capacity_for(sc):
    """Initial agent capacity: twice the initial population (the buffers grow
    on demand)."""
    return 2 * len(sc["agents"]["id"]) + 65536


def summary(pop):
    """A population's counters and the digest of its full state (every live
    agent's fields and genome, qhg_gpu_state_digest)."""
    stats = pop.stats()
    out = {k: int(stats[k]) for k in SUMMARY_KEYS}
    out["digest"] = pop.state_digest()
    return out


def create(sc):
    pop = GpuPopulation.from_scenario(sc, capacity=capacity_for(sc))
    pop.preLoop(0.0)
    return pop


def run_plain(make, n, steps, log=None):
    """make(r) -> scenario dict.  Returns (seconds spent stepping, summaries).
    log(r, seconds) is called after every 32nd scenario."""
    secs = 0.0
    out = []
    for r in range(n):
        if log and r and r % 32 == 0:
            log(r, secs)
        pop = create(make(r))
        pop.sync()
        t0 = time.perf_counter()
        pop.run(steps)
        pop.sync()
        secs += time.perf_counter() - t0
        out.append(summary(pop))
        pop.close()
    return secs, out


class BatchedRun:
    """All n scenarios resident as members of one batch."""

    def __init__(self, make, n, graph=True):
        self.pops = [create(make(r)) for r in range(n)]
        self.batch = GpuBatch()
        for p in self.pops:
            self.batch.add(p)
        self.batch.set_graph(graph)
        self.batch.sync()

    def run(self, steps, tune_steps=2, timing_steps=0):
        """Advance every member by `steps` steps, the autotuner's trial steps
        included.  Returns (seconds, chosen launch shape).  With
        timing_steps, that many of the steps right after the tuning run with
        per-dispatch timing (eagerly, not replayed from the graph), while the
        members hold their early populations; self.timed then holds the
        kernel times and the members' stats before and after them."""
        t0 = time.perf_counter()
        taken = 0
        tuning = self.batch.get_tuning()
        if tune_steps and steps >= 20 * tune_steps:   # the search takes 9 trials
            tuning, taken = self.batch.autotune(tune_steps)
        self.timed = None
        if timing_steps and steps - taken > timing_steps:
            before = [p.stats() for p in self.pops]
            self.batch.set_timing(True)
            self.batch.run(timing_steps)
            self.batch.sync()
            times = self.batch.kernel_times()
            self.batch.set_timing(False)
            self.timed = {"steps": timing_steps, "times": times, "before": before,
                          "after": [p.stats() for p in self.pops]}
            taken += timing_steps
        if steps > taken:
            self.batch.run(steps - taken)
        self.batch.sync()
        return time.perf_counter() - t0, tuning

    def summaries(self):
        return [summary(p) for p in self.pops]

    def device_bytes(self):
        return sum(p.device_bytes() for p in self.pops)

    def close(self):
        self.batch.close()
        for p in self.pops:
            p.close()
        self.pops = []
