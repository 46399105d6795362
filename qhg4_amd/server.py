"""Continuous-batching scenario server on one GPU.

Requests (scenario specs with an arrival time and a horizon in steps) arrive
while others run.  Each arrived request's initial state is prepared on host
worker threads; it is admitted, in arrival order, as soon as a batch slot is
free and its device state fits the HBM budget next to the running members.
Admitted runs become members of one batch (qhg_batch_*): every chunk of steps
is one launch per phase for all members, so the runs overlap on the device.
A member leaves the batch when it reaches its horizon; its final statistics
are read back and its device memory is freed.

While the device works on a chunk, the host admits and creates the next runs
on their own streams; they join the batch between chunks.  Every run ends in
exactly the state it reaches alone (the batch kernels use each member's own
Philox key and step counter).

No counterpart in the reference (one population per process,
QHG4/core/PopLooper.cpp); it serves the continuous-batching pattern.
"""
import collections
import concurrent.futures as cf
import time

import numpy as np

from .ensemble import capacity_for, summary
from .population import GpuBatch, GpuPopulation

# device bytes of a population without genetics: agent buffers (two of 46 B
# per slot) and slot scratch (38 B); per-cell tables at most 340 B per cell
PER_SLOT_BYTES = 130
PER_CELL_BYTES = 340
FIXED_BYTES = 1 << 20
MAX_MEMBERS = 256


def estimate_bytes(sc):
    """Upper estimate of a scenario's device state at admission."""
    return capacity_for(sc) * PER_SLOT_BYTES + int(sc["grid"]["num_cells"]) * PER_CELL_BYTES + FIXED_BYTES


class ScenarioServer:
    def __init__(self, make, budget_bytes=250e9, max_members=MAX_MEMBERS, chunk=16, workers=8, graph=True,
                 tune_every_s=60.0, clock=time.perf_counter):
        """make(request) -> scenario dict (runs on a worker thread)."""
        if not 1 <= max_members <= MAX_MEMBERS:
            raise ValueError("max_members must be in [1, %d]" % MAX_MEMBERS)
        self.make = make
        self.budget = float(budget_bytes)
        self.max_members = max_members
        self.chunk = int(chunk)
        self.workers = workers
        self.graph = graph
        self.tune_every_s = tune_every_s
        self.clock = clock

    def serve(self, requests, log=None):
        """Serve requests (dicts with id, arrival [s from start], horizon)
        in real time.  Returns (results, report)."""
        pending = collections.deque(sorted(requests, key=lambda r: r["arrival"]))
        waiting = collections.deque()       # (request, future scenario)
        active = {}                         # pop -> [request, steps left, admitted]
        joining = []                        # created, not yet members
        results = []
        batch = GpuBatch()
        batch.set_graph(self.graph)
        pool = cf.ThreadPoolExecutor(self.workers)
        t0 = self.clock()
        used = 0
        peak_used = 0
        peak_members = 0
        last_tune = None
        tunings = []
        last_log = t0
        try:
            while pending or waiting or active or joining:
                now = self.clock() - t0
                while pending and pending[0]["arrival"] <= now:
                    r = pending.popleft()
                    waiting.append((r, pool.submit(self.make, r)))
                # members join between chunks
                for pop, r, adm in joining:
                    batch.add(pop)
                    active[pop] = [r, int(r["horizon"]), adm]
                joining = []
                peak_members = max(peak_members, len(active))
                if not active:
                    if not self._admit(waiting, joining, active, used, t0):
                        self._idle(pending, waiting, t0)
                    used = self._used(active, joining)
                    continue
                left = min(e[1] for e in active.values())
                tuned = False
                if (self.tune_every_s is not None and left >= 40 and len(active) >= 8 and
                        (last_tune is None or now - last_tune >= self.tune_every_s)):
                    tuning, taken = batch.autotune(2)
                    tunings.append(dict(tuning, at_s=round(now, 3), members=len(active)))
                    last_tune = now
                    k = taken
                    tuned = True
                else:
                    k = min(self.chunk, left)
                    batch.run(k)
                # while the device works: admit what fits
                self._admit(waiting, joining, active, used, t0)
                batch.sync()
                for e in active.values():
                    e[1] -= k
                done_at = self.clock() - t0
                for pop in [p for p, e in active.items() if e[1] == 0]:
                    r, _, adm = active.pop(pop)
                    batch.remove(pop)
                    s = summary(pop.stats())
                    pop.close()
                    results.append({"id": r["id"], "arrival": r["arrival"], "admitted": adm, "done": done_at,
                                    "latency": done_at - r["arrival"], "summary": s})
                used = self._used(active, joining)
                peak_used = max(peak_used, used)
                if log and (self.clock() - last_log > 10.0 or tuned):
                    last_log = self.clock()
                    log("t=%.1fs done %d running %d queued %d used %.1f GB" %
                        (self.clock() - t0, len(results), len(active), len(waiting), used / 1e9))
        finally:
            for pop, _, _ in joining:
                pop.close()
            for pop in list(active):
                pop.close()
            batch.close()
            pool.shutdown(wait=True)
        t_end = self.clock() - t0
        return results, self._report(results, t_end, peak_used, peak_members, tunings)

    # -- helpers ------------------------------------------------------------
    def _admit(self, waiting, joining, active, used, t0):
        """Create populations for the waiting requests, in arrival order,
        while slots and budget allow.  Returns whether any was created."""
        made = False
        while waiting and waiting[0][1].done() and len(active) + len(joining) < self.max_members:
            r, fut = waiting[0]
            sc = fut.result()
            need = estimate_bytes(sc)
            if used + need > self.budget and (active or joining):
                break
            waiting.popleft()
            pop = GpuPopulation.from_scenario(sc, capacity=capacity_for(sc))
            pop.preLoop(0.0)
            joining.append((pop, r, self.clock() - t0))
            used += pop.device_bytes()
            made = True
        return made

    def _idle(self, pending, waiting, t0):
        if waiting:
            cf.wait([waiting[0][1]], timeout=0.05)
        elif pending:
            time.sleep(max(0.0, min(0.05, pending[0]["arrival"] - (self.clock() - t0))))

    @staticmethod
    def _used(active, joining):
        return sum(p.device_bytes() for p in active) + sum(p.device_bytes() for p, _, _ in joining)

    @staticmethod
    def _report(results, t_end, peak_used, peak_members, tunings):
        lat = np.array([r["latency"] for r in results]) if results else np.zeros(1)
        steps = sum(r["summary"]["agent_steps"] for r in results)
        first = min((r["arrival"] for r in results), default=0.0)
        span = max(t_end - first, 1e-9)
        return {"requests": len(results), "agent_steps": int(steps), "seconds": round(t_end, 3),
                "value": steps / span, "p50_s": float(np.percentile(lat, 50)),
                "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
                "peak_device_bytes": int(peak_used), "peak_members": int(peak_members), "tunings": tunings}
