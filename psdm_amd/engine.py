"""Batched / served RTM engine (psdm_rtm_engine_* in include/psdm_rtm.h).

Many shots migrate concurrently in one device-resident engine: `nslots` shots advance
together, one launch per time step over all of them, with the steps captured in
hipGraphs.  Shots are admitted into free slots when their snapshot store fits the HBM
budget and are stacked into a device-resident FP64 image as they finish.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from ._native import Context

_P, _I, _F, _LL = C.c_void_p, C.c_int, C.c_float, C.c_longlong

SNAP_FORMATS = {"f32": 0, "f16": 1, "bf16": 2}


class EngineCfg(C.Structure):
    """psdm_rtm_engine_cfg."""

    _fields_ = [
        ("nslots", _I), ("ny", _I), ("nx_prop", _I), ("d_vel", _P), ("max_nx", _I), ("max_nt", _I),
        ("dPML", _I), ("fs", _I), ("variant", _I), ("dmodel", _F), ("dt", _F), ("store_snap", _I),
        ("snap_format", _I), ("kahan", _I), ("chunk_steps", _I), ("tile_x", _I), ("stream_w", _I), ("grid", _I),
        ("groups", _I), ("use_graph", _I),
        ("max_inflight", _I), ("hbm_budget", _LL), ("d_psdm_sum", _P), ("d_cov_sum", _P),
        ("steps_per_launch", _I), ("weights", _I),
    ]

WEIGHTS = {"reference": 0, "zero_sum": 1}


class Job(C.Structure):
    """psdm_rtm_job."""

    _fields_ = [
        ("x_i", _I), ("nx_w", _I), ("nt", _I), ("NumRec", _I), ("nxSou", _I), ("nySou", _I),
        ("nxRec", _P), ("nyRec", _P), ("d_wavelet", _P), ("d_data", _P), ("d_image_out", _P),
    ]


def _dptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class Engine:
    """Device-resident multi-shot RTM.  Velocity model, gathers and image are torch
    tensors on the engine's device; `d_vel` is (nx_prop, ny) holding Vpmodel(ny, nx_prop)
    in Fortran order."""

    def __init__(self, ctx: Context, d_vel, ny: int, nx_prop: int, nslots: int, max_nx: int, max_nt: int,
                 dPML: int, fs: int, dmodel: float, dt: float, variant: int = 0, store_snap: int = 1,
                 snap_format: str = "f32", kahan: bool = False, chunk_steps: int = 64, tile_x: int = 32,
                 stream_w: int = 0, grid: int = 0, groups: int = 0, use_graph: bool = True, max_inflight: int = 2, hbm_budget: int = 0,
                 d_psdm_sum=None, d_cov_sum=None, steps_per_launch: int = 1, weights: str = "reference"):
        self.ctx = ctx
        self._L = ctx._L
        self._keep = [d_vel, d_psdm_sum, d_cov_sum]
        cfg = EngineCfg()
        cfg.nslots, cfg.ny, cfg.nx_prop, cfg.d_vel = nslots, ny, nx_prop, d_vel.data_ptr()
        cfg.max_nx, cfg.max_nt, cfg.dPML, cfg.fs, cfg.variant = max_nx, max_nt, dPML, fs, variant
        cfg.dmodel, cfg.dt, cfg.store_snap = dmodel, dt, store_snap
        cfg.snap_format, cfg.kahan = SNAP_FORMATS[snap_format], int(kahan)
        cfg.chunk_steps, cfg.tile_x, cfg.grid, cfg.use_graph = chunk_steps, tile_x, grid, int(use_graph)
        cfg.stream_w, cfg.groups = stream_w, groups
        cfg.max_inflight, cfg.hbm_budget = max_inflight, hbm_budget
        cfg.steps_per_launch = steps_per_launch
        cfg.weights = WEIGHTS[weights]
        cfg.d_psdm_sum = d_psdm_sum.data_ptr() if d_psdm_sum is not None else None
        cfg.d_cov_sum = d_cov_sum.data_ptr() if d_cov_sum is not None else None
        self.cfg = cfg
        h = _P()
        ctx.check(self._L.psdm_rtm_engine_create(ctx.h, C.byref(cfg), C.byref(h)), "psdm_rtm_engine_create")
        self.h = h
        self._jobs = {}
        self.t_submit = {}

    def close(self):
        if getattr(self, "h", None):
            self._L.psdm_rtm_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, x_i, nx_w, nt, nxSou, nySou, nxRec, nyRec, d_wavelet, d_data, d_image_out=None) -> int:
        nxR = np.ascontiguousarray(np.asarray(nxRec, np.int32))
        nyR = np.ascontiguousarray(np.asarray(nyRec, np.int32))
        j = Job()
        j.x_i, j.nx_w, j.nt, j.NumRec = x_i, nx_w, nt, nxR.size
        j.nxSou, j.nySou = nxSou, nySou
        j.nxRec, j.nyRec = nxR.ctypes.data, nyR.ctypes.data
        j.d_wavelet, j.d_data = d_wavelet.data_ptr(), d_data.data_ptr()
        j.d_image_out = d_image_out.data_ptr() if d_image_out is not None else None
        jid = _LL()
        self.ctx.check(self._L.psdm_rtm_engine_submit(self.h, C.byref(j), C.byref(jid)), "psdm_rtm_engine_submit")
        self._jobs[jid.value] = (d_wavelet, d_data, d_image_out)  # keep device buffers alive
        self.t_submit[jid.value] = time.monotonic()
        return jid.value

    def run(self, max_chunks: int = 1):
        na, npend = _I(), _I()
        self.ctx.check(self._L.psdm_rtm_engine_run(self.h, max_chunks, C.byref(na), C.byref(npend)),
                       "psdm_rtm_engine_run")
        return na.value, npend.value

    def poll(self, cap: int = 4096):
        ids = (_LL * cap)()
        ts = (C.c_double * cap)()
        n = _I()
        self.ctx.check(self._L.psdm_rtm_engine_poll(self.h, ids, ts, cap, C.byref(n)), "psdm_rtm_engine_poll")
        out = [(ids[q], ts[q]) for q in range(n.value)]
        for jid, _ in out:
            self._jobs.pop(jid, None)
        return out

    def drain(self):
        self.ctx.check(self._L.psdm_rtm_engine_drain(self.h), "psdm_rtm_engine_drain")

    def timing(self):
        """(mean device ms of one step launch, launches timed), from the chunk events."""
        ms, n = C.c_double(), _LL()
        self.ctx.check(self._L.psdm_rtm_engine_timing(self.h, C.byref(ms), C.byref(n)), "psdm_rtm_engine_timing")
        return ms.value, n.value

    def stats(self):
        a, b, c, d = _LL(), _LL(), _LL(), _LL()
        self.ctx.check(self._L.psdm_rtm_engine_stats(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)),
                       "psdm_rtm_engine_stats")
        return dict(steps_issued=a.value, slot_steps=b.value, hbm_used=c.value, hbm_budget=d.value)

    def memory(self):
        """(fixed device bytes reserved at creation for the slots, snapshot arena bytes)."""
        a, b = _LL(), _LL()
        self.ctx.check(self._L.psdm_rtm_engine_memory(self.h, C.byref(a), C.byref(b)), "psdm_rtm_engine_memory")
        return a.value, b.value
