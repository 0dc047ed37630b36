"""Autotuner of the batched engine's launch shape.

For a workload (model, shot windows, record length) it times short runs of the
engine for each candidate (CPML tile width, interior stream-item width, grid shape,
concurrent slot groups, resident slots) and keeps the one with the lowest expected time per shot.  A trial
admits `nslots` shots with their record length cut to a few chunks (the cost of a step
does not depend on the record length), lets one chunk run (admission and first-step
effects), then times the following chunks, which cover forward and backward steps.
With `nshots` given, a candidate's time per slot-step is scaled by the fill of its
last round of shots (nslots * ceil(nshots / nslots) / nshots).
"""
from __future__ import annotations

import itertools
import time

from .engine import Engine


def _trial(ctx, d_vel, ny, nx_prop, jobs, nslots, max_nx, nt, dPML, fs, dx, dt, cfg, chunk_steps, chunks):
    import torch

    nt_trial = min(nt, (chunks + 1) * chunk_steps // 2 + 1)  # forward and backward inside the trial
    eng = Engine(ctx, d_vel, ny, nx_prop, nslots=nslots, max_nx=max_nx, max_nt=nt, dPML=dPML, fs=fs, dmodel=dx, dt=dt,
                 chunk_steps=chunk_steps, **cfg)
    try:
        for q in range(nslots):
            eng.submit(**dict(jobs[q % len(jobs)], nt=nt_trial))
        eng.run(1)
        ctx.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.run(chunks)
        ctx.sync()
        t1 = time.perf_counter()
    finally:
        eng.close()
    return (t1 - t0) / (chunks * chunk_steps * nslots)


def tune(ctx, d_vel, ny, nx_prop, jobs, max_nx, nt, dPML, fs, dx, dt, snap_format="f32", store_snap=1, kahan=False,
         slots=(16, 24, 32), tiles=(16, 32), stream_ws=(4, 8), grids=(0,), groups=(2,), chunk_steps=32, chunks=5,
         nshots=None, log=None):
    """jobs: list of Engine.submit keyword dicts.  Returns (best config dict, trials)."""
    trials = []
    for nslots, tile_x, sw, grid, ng in itertools.product(slots, tiles, stream_ws, grids, groups):
        cfg = dict(tile_x=tile_x, stream_w=sw, grid=grid, groups=ng, snap_format=snap_format, store_snap=store_snap,
                   kahan=kahan)
        try:
            t = _trial(ctx, d_vel, ny, nx_prop, jobs, nslots, max_nx, nt, dPML, fs, dx, dt, cfg, chunk_steps, chunks)
        except Exception as exc:  # e.g. a slot count whose snapshot stores exceed HBM
            if log:
                log(f"[autotune] slots={nslots} tile_x={tile_x} stream_w={sw} grid={grid} groups={ng}: skipped ({exc})")
            continue
        fill = nslots * -(-nshots // nslots) / nshots if nshots else 1.0
        trials.append(dict(nslots=nslots, tile_x=tile_x, stream_w=sw, grid=grid, groups=ng, us_per_slot_step=1e6 * t,
                           score=1e6 * t * fill))
        if log:
            log(f"[autotune] slots={nslots} tile_x={tile_x} stream_w={sw} grid={grid} groups={ng}: {1e6 * t:.3f} us per slot-step"
                f" (x{fill:.3f} last-round fill)")
    best = min(trials, key=lambda r: r["score"])
    return best, trials
