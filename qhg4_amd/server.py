"""Continuous-batching scenario server on one GPU.

Requests (scenario specs with an arrival time and a horizon in steps) arrive
while others run.  Each arrived request's initial state is prepared on host
worker threads; it is admitted, in arrival order, as soon as a batch slot is
free and its device state fits the HBM budget next to the running members.
Admitted runs become members of one batch (qhg_batch_*): every chunk of steps
is one launch per phase for all members, so the runs overlap on the device.
A member leaves the batch when it reaches its horizon; its final statistics
are read back and its device memory is freed.

A request that fails (its scenario or population cannot be built, uploaded or
joined; its device step reports an error; its buffers cannot grow; its
result cannot be read) is reported with an "error" and its slot and memory
are freed; the other requests go on.

Admission reserves the request's estimated device memory; a worker thread
then builds its initial state, creates the population on its own stream and
uploads it while the device works on the batch; it joins between chunks.  Every run ends in
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
# (the population reports its exact figure once it exists)
PER_SLOT_BYTES = 130
PER_CELL_BYTES = 340
FIXED_BYTES = 1 << 20
MAX_MEMBERS = 256


def estimate_bytes(sc):
    """Upper estimate of a scenario's device state at admission."""
    return capacity_for(sc) * PER_SLOT_BYTES + int(sc["grid"]["num_cells"]) * PER_CELL_BYTES + FIXED_BYTES


def estimate_request(req):
    """Device bytes of a C5-style request (its "n0" agents on a level-"level"
    icosahedral grid) before its state exists."""
    ncells = 10 * 4 ** int(req["level"]) + 2
    return (2 * int(req["n0"]) + 65536) * PER_SLOT_BYTES + ncells * PER_CELL_BYTES + FIXED_BYTES


class ScenarioServer:
    def __init__(self, make, estimate=estimate_request, budget_bytes=250e9, max_members=MAX_MEMBERS, chunk=16,
                 workers=8, graph=True, tune_every_s=60.0, clock=time.perf_counter):
        """make(request) -> scenario dict; estimate(request) -> device bytes
        its state will take (reserved at admission).  Both run before the
        request's state exists; make runs on a worker thread."""
        if not 1 <= max_members <= MAX_MEMBERS:
            raise ValueError("max_members must be in [1, %d]" % MAX_MEMBERS)
        self.make = make
        self.estimate = estimate
        self.budget = float(budget_bytes)
        self.max_members = max_members
        self.chunk = int(chunk)
        self.workers = workers
        self.graph = graph
        self.tune_every_s = tune_every_s
        self.clock = clock

    def _prepare(self, req):
        """Worker thread: the request's scenario, its population on the
        device (uploaded, pre-loop done), ready to join the batch."""
        sc = self.make(req)
        pop = GpuPopulation.from_scenario(sc, capacity=capacity_for(sc))
        pop.preLoop(0.0)
        return pop

    def serve(self, requests, log=None):
        """Serve requests (dicts with id, arrival [s from start], horizon)
        in real time.  Returns (results, report)."""
        pending = collections.deque(sorted(requests, key=lambda r: r["arrival"]))
        waiting = collections.deque()       # arrived, not yet admitted
        preparing = []                      # [request, future population, reserved bytes, admitted at]
        active = {}                         # pop -> [request, steps left, admitted at, joined at]
        results = []
        batch = GpuBatch()
        batch.set_graph(self.graph)
        pool = cf.ThreadPoolExecutor(self.workers)
        t0 = self.clock()
        used = 0                            # device bytes of the members
        peak_used = 0
        peak_members = 0
        last_tune = None
        tunings = []
        last_log = t0
        try:
            while pending or waiting or preparing or active:
                now = self.clock() - t0
                while pending and pending[0]["arrival"] <= now:
                    waiting.append(pending.popleft())
                # admission in arrival order while slots and the budget allow
                reserved = sum(j[2] for j in preparing)
                while waiting and len(active) + len(preparing) < self.max_members:
                    need = int(self.estimate(waiting[0]))
                    if used + reserved + need > self.budget and (active or preparing):
                        break
                    r = waiting.popleft()
                    preparing.append([r, pool.submit(self._prepare, r), need, self.clock() - t0])
                    reserved += need
                # prepared populations join between chunks
                for job in [j for j in preparing if j[1].done()]:
                    preparing.remove(job)
                    pop = None
                    try:
                        pop = job[1].result()
                        batch.add(pop)
                    except Exception as e:  # noqa: BLE001 -- one request's failure
                        if pop is not None:
                            pop.close()
                        results.append(self._failed(job[0], job[3], None, self.clock() - t0, e))
                        continue
                    active[pop] = [job[0], int(job[0]["horizon"]), job[3], self.clock() - t0]
                    used += pop.device_bytes()
                peak_members = max(peak_members, len(active))
                peak_used = max(peak_used, used + sum(j[2] for j in preparing))
                if not active:
                    self._idle(pending, preparing, t0)
                    continue
                left = min(e[1] for e in active.values())
                tuned = False
                if (self.tune_every_s is not None and left >= 40 and len(active) >= 8 and
                        (last_tune is None or now - last_tune >= self.tune_every_s)):
                    tuning, k = batch.autotune(2)
                    tunings.append(dict(tuning, at_s=round(now, 3), members=len(active)))
                    last_tune = now
                    tuned = True
                else:
                    k = min(self.chunk, left)
                    batch.run(k)
                batch.sync()
                done_at = self.clock() - t0
                for pop in batch.take_failed():      # left the batch on an error
                    r, _, adm, joined = active.pop(pop)
                    used -= pop.device_bytes()
                    err = "step failed (code %d)" % pop.stats()["host_error"]
                    pop.close()
                    results.append(self._failed(r, adm, joined, done_at, err))
                for e in active.values():
                    e[1] -= k
                for pop in [p for p, e in active.items() if e[1] == 0]:
                    r, _, adm, joined = active.pop(pop)
                    used -= pop.device_bytes()
                    try:
                        batch.remove(pop)
                        s = summary(pop)
                    except Exception as e:  # noqa: BLE001 -- one request's failure
                        pop.close()
                        results.append(self._failed(r, adm, joined, done_at, e))
                        continue
                    pop.close()
                    results.append({"id": r["id"], "arrival": r["arrival"], "admitted": adm, "joined": joined,
                                    "done": done_at, "latency": done_at - r["arrival"], "summary": s})
                if log and (self.clock() - last_log > 10.0 or tuned):
                    last_log = self.clock()
                    log("t=%.1fs done %d running %d preparing %d queued %d used %.1f GB" %
                        (self.clock() - t0, len(results), len(active), len(preparing), len(waiting), used / 1e9))
        finally:
            for job in preparing:
                try:
                    job[1].result().close()
                except Exception:
                    pass
            for pop in list(active):
                pop.close()
            batch.close()
            pool.shutdown(wait=True)
        t_end = self.clock() - t0
        return results, self._report(results, t_end, peak_used, peak_members, tunings)

    # -- helpers ------------------------------------------------------------
    @staticmethod
    def _failed(r, admitted, joined, done, err):
        return {"id": r["id"], "arrival": r["arrival"], "admitted": admitted, "joined": joined, "done": done,
                "latency": done - r["arrival"], "summary": None, "error": str(err)}

    def _idle(self, pending, preparing, t0):
        if preparing:
            cf.wait([j[1] for j in preparing], timeout=0.05, return_when=cf.FIRST_COMPLETED)
        elif pending:
            time.sleep(max(0.0, min(0.05, pending[0]["arrival"] - (self.clock() - t0))))

    @staticmethod
    def _report(results, t_end, peak_used, peak_members, tunings):
        ok = [r for r in results if r["summary"] is not None]
        lat = np.array([r["latency"] for r in ok]) if ok else np.zeros(1)
        steps = sum(r["summary"]["agent_steps"] for r in ok)
        first = min((r["arrival"] for r in results), default=0.0)
        span = max(t_end - first, 1e-9)
        return {"requests": len(results), "failed": len(results) - len(ok), "agent_steps": int(steps),
                "seconds": round(t_end, 3),
                "value": steps / span, "p50_s": float(np.percentile(lat, 50)),
                "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
                "peak_device_bytes": int(peak_used), "peak_members": int(peak_members), "tunings": tunings}
