"""Served RTM: a continuous-batching request loop over device-resident engines.

Requests are shot gathers of uneven aperture and record length that arrive over time
(Poisson arrivals in the benchmark).  Each request names the engine that migrates it --
one engine per migration velocity model, the model and the partial image staying on the
device -- and is handed to that engine as soon as it arrives.  The engine admits it into
a free slot at the next chunk boundary once its snapshot store fits the engine's HBM
budget, so requests that arrive while others are on the GPU join the running launch
(continuous batching, not a queue in front of a one-shot solver).  The loop issues one
chunk per engine in turn, polls completions and records, per request, the latency from
arrival to the image being stacked, plus the HBM occupancy of the snapshot stores.

The request path of the reference is one shot at a time per MPI rank
(calculate_PSDM_RTM.f90:377-497); the image of every request here is the one-shot
solver's image (bit for bit in FP32), whatever else shares the device.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    arrival: float   # seconds after the stream starts
    engine: str      # key of the engine (velocity model) that migrates the gather
    job: dict        # Engine.submit keyword arguments (device buffers stay owned by the caller)
    kind: str = ""
    meta: dict = field(default_factory=dict)


def poisson_arrivals(rate: float, n: int, seed: int = 0) -> np.ndarray:
    """Arrival times (s) of n requests of a Poisson process with `rate` per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def latency_stats(lat) -> dict:
    lat = np.asarray(lat, np.float64)
    if lat.size == 0:
        return dict(n=0, p50_s=None, p99_s=None, mean_s=None, max_s=None)
    return dict(n=int(lat.size), p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_s=float(lat.mean()), max_s=float(lat.max()))


class Server:
    """engines: {name: psdm_amd.engine.Engine}; fixed_bytes: {name: device bytes the
    engine holds besides its snapshot stores (wavefields, tables, images)}."""

    def __init__(self, engines: dict, fixed_bytes: dict | None = None):
        if not engines:
            raise ValueError("a server needs at least one engine")
        self.engines = engines
        self.fixed_bytes = fixed_bytes or {k: 0 for k in engines}

    def run(self, requests, chunks_per_turn: int = 1, log=None) -> dict:
        """Serves `requests` (arrival order is by Request.arrival) to completion."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            if r.engine not in self.engines:
                raise KeyError(f"request {r.rid} names unknown engine {r.engine!r}")
        by_job = {}
        submitted = {}
        done = {}
        occ_t, occ_b = [], []
        budget = sum(e.stats()["hbm_budget"] for e in self.engines.values())
        fixed = sum(self.fixed_bytes.values())
        active = {k: False for k in self.engines}
        t0 = time.monotonic()
        nxt = 0
        while len(done) < len(reqs):
            now = time.monotonic() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                r = reqs[nxt]
                jid = self.engines[r.engine].submit(**r.job)
                by_job[(r.engine, jid)] = r
                submitted[r.rid] = time.monotonic() - t0
                active[r.engine] = True
                nxt += 1
            if not any(active.values()):
                # idle until the next arrival
                time.sleep(max(0.0, min(reqs[nxt].arrival - (time.monotonic() - t0), 0.01)))
                continue
            for name, eng in self.engines.items():
                if not active[name]:
                    continue
                n_act, n_pend = eng.run(chunks_per_turn)
                for jid, t_done in eng.poll():
                    r = by_job.pop((name, jid))
                    done[r.rid] = t_done - t0
                if n_act == 0 and n_pend == 0 and not any(k[0] == name for k in by_job):
                    active[name] = False
            occ_t.append(time.monotonic() - t0)
            occ_b.append(fixed + sum(e.stats()["hbm_used"] for e in self.engines.values()))
        t_end = max(done.values()) if done else 0.0
        lat = {r.rid: done[r.rid] - r.arrival for r in reqs}
        kinds = sorted({r.kind for r in reqs})
        out = dict(
            n=len(reqs), makespan_s=t_end, shots_per_s=len(reqs) / t_end if t_end > 0 else None,
            latency=latency_stats(list(lat.values())),
            by_kind={k: latency_stats([lat[r.rid] for r in reqs if r.kind == k]) for k in kinds},
            hbm={"budget_bytes": int(budget + fixed), "mean_bytes": float(np.mean(occ_b)) if occ_b else 0.0,
                 "max_bytes": int(max(occ_b)) if occ_b else 0},
            requests=[dict(rid=r.rid, kind=r.kind, arrival=r.arrival, submitted=submitted[r.rid], done=done[r.rid],
                           latency=lat[r.rid]) for r in reqs])
        if log:
            L = out["latency"]
            log(f"[serve] {out['n']} requests in {t_end:.2f} s: p50 {L['p50_s']:.3f} s, p99 {L['p99_s']:.3f} s")
        return out
