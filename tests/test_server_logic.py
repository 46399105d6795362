"""The scenario server's host logic (admission order, slot and memory
limits, horizons), with the device stepping replaced by a stand-in that
counts steps: no GPU needed."""
import time

import numpy as np
import pytest

from qhg4_amd import server


class FakePop:
    live = []

    def __init__(self, sc):
        self.sc = sc
        self.steps = 0
        self.bytes = server.estimate_bytes(sc) - 4096
        FakePop.live.append(self)

    @classmethod
    def from_scenario(cls, sc, capacity=0):
        return cls(sc)

    def preLoop(self, t=0.0):
        pass

    def device_bytes(self):
        return self.bytes

    def stats(self):
        n = len(self.sc["agents"]["id"])
        return {k: 0 for k in server.summary.__globals__["SUMMARY_KEYS"]} | {"step": self.steps,
                                                                              "agent_steps": n * self.steps,
                                                                              "host_error": self.sc.get("err", 0)}

    def state_digest(self):
        if self.sc.get("bad_readback"):
            raise RuntimeError("read-back failed")
        return "%d:%d" % (len(self.sc["agents"]["id"]), self.steps)

    def close(self):
        if self in FakePop.live:
            FakePop.live.remove(self)


class FakeBatch:
    def __init__(self):
        self.members = []
        self.max_seen = 0
        self.tuned = 0
        self.failed = []

    def set_graph(self, on):
        pass

    def add(self, p):
        self.members.append(p)
        self.max_seen = max(self.max_seen, len(self.members))

    def remove(self, p):
        self.members.remove(p)

    def run(self, k):
        for p in list(self.members):
            fail_at = p.sc.get("fail_at")
            if fail_at is not None and p.steps + k > fail_at:
                # the device step of this member fails: it leaves the batch
                p.steps = fail_at
                p.sc["err"] = -3
                self.members.remove(p)
                self.failed.append(p)
                continue
            p.steps += k

    def take_failed(self):
        out, self.failed = self.failed, []
        return out

    def autotune(self, trial):
        self.tuned += 1
        self.run(9 * trial)
        return {"group_agents": 64.0, "blocks_per_cu": 16}, 9 * trial

    def sync(self):
        time.sleep(2e-4)     # a device wait releases the GIL to the preparing threads

    def close(self):
        pass


class Clock:
    """Virtual time: every call advances 1 ms."""

    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-3
        return self.t


def fake_scenario(req):
    n = req["n0"]
    return {"grid": {"num_cells": 1000}, "agents": {"id": np.arange(n)}}


def fake_estimate(req):
    return server.estimate_bytes(fake_scenario(req))


@pytest.fixture
def fakes(monkeypatch):
    made = []

    def batch():
        b = FakeBatch()
        made.append(b)
        return b
    monkeypatch.setattr(server, "GpuPopulation", FakePop)
    monkeypatch.setattr(server, "GpuBatch", batch)
    FakePop.live = []
    return made


def requests(n, seed=1):
    rng = np.random.default_rng(seed)
    arr = np.cumsum(rng.exponential(0.005, n))
    return [{"id": i, "arrival": float(arr[i]), "n0": int(rng.integers(1000, 200000)),
             "horizon": int(rng.integers(1, 120))} for i in range(n)]


@pytest.mark.parametrize("max_members,budget_mb,chunk", [(4, 1e6, 16), (256, 60.0, 8), (3, 40.0, 1)])
def test_every_request_runs_exactly_its_horizon(fakes, max_members, budget_mb, chunk):
    reqs = requests(60)
    srv = server.ScenarioServer(fake_scenario, estimate=fake_estimate, budget_bytes=budget_mb * 1e6,
                                max_members=max_members, chunk=chunk, workers=2, tune_every_s=None, clock=Clock())
    res, rep = srv.serve(reqs)
    assert sorted(r["id"] for r in res) == list(range(60))
    by = {r["id"]: r for r in res}
    for q in reqs:
        r = by[q["id"]]
        assert r["summary"]["step"] == q["horizon"]
        assert r["admitted"] >= q["arrival"]
        assert r["done"] >= r["admitted"]
    # first come, first admitted (joining the batch follows preparation)
    adm = [by[q["id"]]["admitted"] for q in sorted(reqs, key=lambda q: q["arrival"])]
    assert adm == sorted(adm)
    assert all(by[q["id"]]["joined"] >= by[q["id"]]["admitted"] for q in reqs)
    assert fakes[0].max_seen <= max_members
    one = max(server.estimate_bytes(fake_scenario(q)) for q in reqs)
    assert rep["peak_device_bytes"] <= max(budget_mb * 1e6, one)
    assert rep["p99_s"] >= rep["p50_s"] > 0
    assert not FakePop.live                     # every population was freed


def test_autotune_never_overshoots_a_horizon(fakes):
    reqs = [dict(q, arrival=0.0, horizon=1000 + 7 * q["id"]) for q in requests(12)]
    srv = server.ScenarioServer(fake_scenario, estimate=fake_estimate, budget_bytes=1e12, max_members=16, chunk=16,
                                workers=2, tune_every_s=0.0, clock=Clock())
    res, rep = srv.serve(reqs)
    assert fakes[0].tuned >= 1 and rep["tunings"]
    for r in res:
        assert r["summary"]["step"] == reqs[r["id"]]["horizon"]


def test_server_rejects_bad_member_limit():
    with pytest.raises(ValueError):
        server.ScenarioServer(fake_scenario, max_members=300)


def failing_scenario(req):
    if req.get("bad_make"):
        raise ValueError("no such grid")
    sc = fake_scenario(req)
    for k in ("fail_at", "bad_readback"):
        if k in req:
            sc[k] = req[k]
    return sc


@pytest.mark.parametrize("chunk", [1, 8])
def test_a_failing_request_is_evicted_alone(fakes, chunk):
    """A request whose scenario cannot be built, one whose device step fails
    midway and one whose result cannot be read are reported as failed; every
    other request still runs its full horizon, and nothing stays allocated."""
    reqs = requests(40, seed=3)
    reqs[3]["bad_make"] = True
    reqs[11]["fail_at"] = max(1, reqs[11]["horizon"] // 2)
    reqs[11]["horizon"] = max(reqs[11]["horizon"], 10)
    reqs[20]["bad_readback"] = True
    srv = server.ScenarioServer(failing_scenario, estimate=fake_estimate, budget_bytes=80e6, max_members=6,
                                chunk=chunk, workers=2, tune_every_s=None, clock=Clock())
    res, rep = srv.serve(reqs)
    assert sorted(r["id"] for r in res) == list(range(40))
    by = {r["id"]: r for r in res}
    bad = {3, 11, 20}
    for q in reqs:
        r = by[q["id"]]
        if q["id"] in bad:
            assert r["summary"] is None and r["error"], q["id"]
        else:
            assert "error" not in r and r["summary"]["step"] == q["horizon"], q["id"]
    assert "no such grid" in by[3]["error"] and "read-back" in by[20]["error"]
    assert rep["failed"] == 3 and rep["requests"] == 40
    assert not FakePop.live
    assert not fakes[0].members
