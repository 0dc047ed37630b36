"""Autotuner of the batched engine's launch shape.

For a workload (model, shot windows, record length) it times short runs of the engine
and keeps the configuration with the lowest expected time per shot.  A trial admits
shots with their record length cut to a few chunks (the cost of a step does not depend on
the record length), lets one chunk run (admission and first-step effects), then times the
following chunks, which cover forward and backward steps.

The search has two stages.  (1) The launch shape -- CPML tile width, interior
stream-item width, concurrent slot groups, and time steps per launch (1, or 2 with
temporal blocking) -- is chosen at the first slot count.  (2) With that shape, the number of resident slots S is chosen by
the time the engine would take for the workload's `nshots` shots of equal record length:
floor(nshots / S) rounds with all S slots busy plus, when S does not divide nshots, one
round with the remainder busy, each timed at its own occupancy (a step with few busy
slots does not cost proportionally less: its launches fill less of the chip).  Without
`nshots` every candidate is scored by its time per busy slot-step.
"""
from __future__ import annotations

import itertools
import time

from .engine import Engine


def _trial(ctx, d_vel, ny, nx_prop, jobs, nslots, busy, max_nx, nt, dPML, fs, dx, dt, cfg, chunk_steps, chunks,
           hbm_budget=0):
    """Seconds per engine step with `busy` of `nslots` slots occupied."""
    import torch

    cfg = dict(cfg)
    spl = cfg.pop("steps_per_launch", 1)
    if spl != 1:
        cfg["steps_per_launch"] = spl
    nt_trial = min(nt, (chunks + 1) * chunk_steps * spl // 2 + 1)  # forward and backward inside the trial
    eng = Engine(ctx, d_vel, ny, nx_prop, nslots=nslots, max_nx=max_nx, max_nt=nt, dPML=dPML, fs=fs, dmodel=dx, dt=dt,
                 chunk_steps=chunk_steps, hbm_budget=hbm_budget, **cfg)
    try:
        for q in range(busy):
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
    return (t1 - t0) / (chunks * chunk_steps * spl)


def slot_candidates(nshots, base=(16, 24, 32), lo=8, hi=32):
    """The base slot counts plus the ones that split `nshots` into 1..4 equal rounds."""
    c = set(base)
    if nshots:
        for r in range(1, 5):
            s = -(-nshots // r)
            if lo <= s <= hi:
                c.add(s)
    return sorted(c)


def round_time(t_step, nshots, nslots, t_part=None):
    """Engine steps' time for nshots equal shots through nslots slots, per shot-step."""
    full, rem = divmod(nshots, nslots)
    return (full * t_step + (t_part if rem else 0.0)) / nshots


def tune(ctx, d_vel, ny, nx_prop, jobs, max_nx, nt, dPML, fs, dx, dt, snap_format="f32", store_snap=1, kahan=False,
         slots=None, tiles=(16, 32), stream_ws=(4, 8), grids=(0,), groups=(2,), steps_per_launch=(1,),
         tb_stream_ws=(8, 16), chunk_steps=32, chunks=5, nshots=None, hbm_budget=0, log=None):
    """jobs: list of Engine.submit keyword dicts.  hbm_budget: the trial engines' snapshot
    arena (0 = the engine's default, nslots x the largest shot at `nt`; give the workload's
    budget where that would not fit in HBM).  Returns (best config dict, trials)."""
    slots = list(slots) if slots else slot_candidates(nshots)
    say = log or (lambda *a: None)
    trials = []
    base = dict(snap_format=snap_format, store_snap=store_snap, kahan=kahan)

    def run(nslots, busy, cfg):
        try:
            return _trial(ctx, d_vel, ny, nx_prop, jobs, nslots, busy, max_nx, nt, dPML, fs, dx, dt, dict(base, **cfg),
                          chunk_steps, chunks, hbm_budget)
        except Exception as exc:  # e.g. a slot count whose snapshot stores exceed HBM
            say(f"[autotune] slots={nslots} busy={busy} {cfg}: skipped ({exc})")
            return None

    # (1) launch shape at the first slot count, every slot busy.  Two steps per launch have
    # fixed edge tiles (tile_x does not apply) and wider interior items (each computes its
    # first step over 3 extra columns on either side)
    s0 = 16 if 16 in slots else slots[0]
    shapes = []
    cands = []
    for spl in steps_per_launch:
        if spl == 2:
            cands += [(tiles[0], sw, grid, ng, 2) for sw, grid, ng in itertools.product(tb_stream_ws, grids, groups)]
        else:
            cands += [(t, sw, grid, ng, spl) for t, sw, grid, ng in itertools.product(tiles, stream_ws, grids, groups)]
    for tile_x, sw, grid, ng, spl in cands:
        cfg = dict(tile_x=tile_x, stream_w=sw, grid=grid, groups=ng, steps_per_launch=spl)
        t = run(s0, s0, cfg)
        if t is None:
            continue
        shapes.append((t, cfg))
        trials.append(dict(nslots=s0, busy=s0, **cfg, us_per_slot_step=1e6 * t / s0))
        say(f"[autotune] slots={s0} {cfg}: {1e6 * t / s0:.3f} us per slot-step")
    if not shapes:
        raise RuntimeError("autotune: no launch shape could be run")
    shape = min(shapes, key=lambda r: r[0])[1]

    # (2) resident slots with that shape, scored over the workload's rounds
    best = None
    for S in slots:
        t = next((t for t, c in shapes if c == shape), None) if S == s0 else run(S, S, shape)
        if t is None:
            continue
        rec = dict(nslots=S, **shape, us_per_slot_step=1e6 * t / S)
        if nshots:
            rem = nshots % S
            tp = run(S, rem, shape) if rem else None
            if rem and tp is None:
                continue
            rec["score"] = 1e6 * round_time(t, nshots, S, tp)
            rec["us_per_step_partial_round"] = 1e6 * tp if tp else None
        else:
            rec["score"] = rec["us_per_slot_step"]
        trials.append(rec)
        say(f"[autotune] slots={S}: {rec['us_per_slot_step']:.3f} us per slot-step, "
            f"{rec['score']:.3f} us per shot-step over {nshots or 'full'} shots")
        if best is None or rec["score"] < best["score"]:
            best = rec
    if best is None:
        raise RuntimeError("autotune: no slot count could be run")
    return best, trials


def tune_plain(ctx, probe, tiles=(8, 16, 32), widths=(1, 2, 4, 8), timing_every=2, log=None):
    """Launch shape of the one-shot solver (CPML tile width, interior columns per stream
    item) for one window: `probe()` runs a short shot through the device entry points
    (forward + backward on that window); each shape is scored by the mean duration of its
    step kernel (start/stop stamps on sampled dispatches).  Leaves the best shape set on
    the context and returns (best, trials)."""
    trials = []
    for tx, sw in itertools.product(tiles, widths):
        ctx.set_launch(tile_x=tx)
        ctx.set_stream_width(sw)
        probe()  # allocations and first-touch of this shape's buffers
        ctx.sync()
        ctx.set_timing(timing_every)
        probe()
        ctx.sync()
        n, ms = ctx.step_stats()
        ctx.set_timing(0)
        trials.append(dict(tile_x=tx, stream_w=sw, kernel_ms=round(ms, 5), sampled=n))
        if log:
            log(f"[tune plain] tile_x={tx} stream_w={sw}: {ms * 1e3:.2f} us per step kernel")
    best = min(trials, key=lambda t: t["kernel_ms"])
    ctx.set_launch(tile_x=best["tile_x"])
    ctx.set_stream_width(best["stream_w"])
    return best, trials
