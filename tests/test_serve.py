"""CPU tests of the serving loop (psdm_amd.serve) with a stand-in engine: arrival
process, admission order, latency bookkeeping and idle handling."""
import time

import numpy as np
import pytest

from psdm_amd.serve import Request, Server, latency_stats, poisson_arrivals


def test_poisson_arrivals_seeded_and_rate():
    a = poisson_arrivals(50.0, 4000, seed=3)
    b = poisson_arrivals(50.0, 4000, seed=3)
    np.testing.assert_array_equal(a, b)
    assert np.all(np.diff(a) > 0)
    gaps = np.diff(np.concatenate([[0.0], a]))
    assert abs(gaps.mean() - 1 / 50.0) < 0.1 / 50.0  # exponential gaps with mean 1/rate


def test_latency_stats():
    s = latency_stats(np.arange(1, 101, dtype=float))
    assert s["n"] == 100 and s["p50_s"] == pytest.approx(50.5) and s["p99_s"] == pytest.approx(99.01)
    assert latency_stats([])["p50_s"] is None


class FakeEngine:
    """Completes a job after `work` run() chunks, at most `slots` at a time, FIFO."""

    def __init__(self, slots=2, budget=1000):
        self.slots, self.budget = slots, budget
        self.queue, self.active, self.ready, self.log = [], {}, [], []
        self.next_id = 0

    def submit(self, work, tag):
        jid = self.next_id
        self.next_id += 1
        self.queue.append((jid, work, tag))
        self.log.append(("submit", tag))
        return jid

    def run(self, n):
        for _ in range(n):
            while self.queue and len(self.active) < self.slots:
                jid, work, tag = self.queue.pop(0)
                self.active[jid] = [work, tag]
            if not self.active:
                break
            for jid in list(self.active):
                self.active[jid][0] -= 1
                if self.active[jid][0] == 0:
                    self.ready.append((jid, time.monotonic()))
                    self.log.append(("done", self.active.pop(jid)[1]))
        return len(self.active), len(self.queue)

    def poll(self):
        out, self.ready = self.ready, []
        return out

    def stats(self):
        return dict(hbm_used=100 * len(self.active), hbm_budget=self.budget, steps_issued=0, slot_steps=0)


def test_server_serves_every_request_in_arrival_order_with_overlap():
    eng = FakeEngine(slots=2)
    arr = [0.0, 0.0, 0.0, 0.02, 0.05]
    reqs = [Request(rid=q, arrival=t, engine="m", job=dict(work=3, tag=q), kind="mcs") for q, t in enumerate(arr)]
    out = Server({"m": eng}, fixed_bytes={"m": 7}).run(reqs)
    assert out["n"] == 5 and out["latency"]["n"] == 5
    subs = [t for k, t in eng.log if k == "submit"]
    assert subs == [0, 1, 2, 3, 4]
    # two slots: requests 0 and 1 run together (overlap), 2 waits for a free slot
    done = [t for k, t in eng.log if k == "done"]
    assert set(done[:2]) == {0, 1}
    for r in out["requests"]:
        assert r["done"] >= r["arrival"] and r["latency"] == pytest.approx(r["done"] - r["arrival"])
    assert out["hbm"]["budget_bytes"] == 1007 and out["hbm"]["max_bytes"] <= 207


def test_server_idles_until_late_arrival_and_routes_engines():
    a, b = FakeEngine(), FakeEngine()
    reqs = [Request(0, 0.0, "a", dict(work=1, tag="x"), kind="mcs"),
            Request(1, 0.08, "b", dict(work=2, tag="y"), kind="was")]
    t0 = time.monotonic()
    out = Server({"a": a, "b": b}).run(reqs)
    assert time.monotonic() - t0 >= 0.08
    assert [t for k, t in a.log if k == "done"] == ["x"] and [t for k, t in b.log if k == "done"] == ["y"]
    assert set(out["by_kind"]) == {"mcs", "was"}
    assert out["requests"][1]["submitted"] >= 0.08


def test_server_rejects_unknown_engine():
    with pytest.raises(KeyError):
        Server({"a": FakeEngine()}).run([Request(0, 0.0, "zz", dict(work=1, tag=0))])
