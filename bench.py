"""RTM line migration benchmark on one MI355X.

Workload (default): the 2001 x 601 @ 10 m marine line -- 64 shots at 25 m spacing,
480 channels at 12.5 m from 150 m offset, dt 0.8 ms, nt 6000, CPML 40 cells, each shot
migrated in its 715-column aperture window (forward with FP32 snapshots every step in
HBM, back-propagation of the time-reversed gather, zero-lag imaging, FP64 cross-shot
stack).  One "step" = one migrated shot; the timed region migrates K shots with the
velocity model, the gathers and the image already resident in HBM.

Prints one JSON line: shots/s (value), Gpt-updates/s, the step kernel's HBM roofline
(algorithmic bytes / mean kernel duration from HIP events on the library stream) and
the CPU baseline (reference Fortran solvers, N processes) on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """(cores to use, this host's physical cores, logical CPUs): every CPU this process may
    run on, at most 16 (the host cores allotted to one GPU of the box)."""
    avail = len(os.sched_getaffinity(0))
    phys = set()
    try:
        pid = core = None
        for line in open("/proc/cpuinfo"):
            if line.startswith("physical id"):
                pid = line.split(":")[1].strip()
            elif line.startswith("core id"):
                core = line.split(":")[1].strip()
                phys.add((pid, core))
    except OSError:
        pass
    return min(avail, 16), len(phys) or None, os.cpu_count()


def cpu_leg(ny, nx_w, dPML, dx, dt, nprocs, nt_sample, nt_full):
    from oracle import cpu_baseline

    r = cpu_baseline.run(ny, nx_w, nt_sample, dPML, dx, dt, nprocs)
    dn = dPML + 6
    upd_shot = 2 * (ny + 2 * dn) * (nx_w + 2 * dn) * (nt_full - 1)
    _, phys, logical = host_cores()
    return {
        "value": round(r["gupd_per_s"], 4),
        "unit": "Gpt-updates/s",
        "shots_per_s": round(r["gupd_per_s"] * 1e9 / upd_shot, 6),
        "cores": r["cores"],
        "kind": r["kind"],
        "host": {"cpus_allowed": len(os.sched_getaffinity(0)), "physical_cores": phys, "logical_cpus": logical},
        "sample": (f"{r['cores']} independent single-threaded processes (one per core, as mpirun ranks), each "
                   f"forward+backward of one {ny}x{nx_w} window (dPML {dPML}) over {nt_sample} of the {nt_full} "
                   f"steps; rate = updates / slowest rank time ({r['rank_s']:.1f} s)"),
    }


def accuracy_check(ctx):
    """Single-shot 401 x 201 @ 12.5 m workload cut to 600 steps: the gather is modelled on
    the true model on the GPU, migrated on the GPU (plain form), and the image and the
    forward traces are compared with the FP64 oracle's, kept as a fixture in
    tests/golden (accuracy_single401_fp64.npz)."""
    import numpy as np

    from psdm_amd import solvers as S
    from psdm_amd.workloads import line_workload
    from tests.cases import rel_l2

    fx = np.load(os.path.join(ROOT, "tests", "golden", "accuracy_single401_fp64.npz"))
    W = line_workload("single401")
    m, sh = W["model"], W["shots"][0]
    nt, dt, dPML = int(fx["nt"]), W["dt"], W["dPML"]
    ctmp = np.asfortranarray(m.migration[:, sh.x_i - 1:sh.x_f])
    ctrue = np.asfortranarray(m.true[:, sh.x_i - 1:sh.x_f])
    src = S.ricker_mcs(nt, dt, W["f0"], 1.2 / W["f0"]).reshape(nt, 1)
    nxS, nyS = [sh.nxSou - sh.x_i + 1], [sh.nySou]
    nxR, nyR = sh.nxRec - sh.x_i + 1, sh.nyRec
    ny, nx = ctmp.shape
    par = S.ParFile(dPML=dPML, fs=0)
    NumRec = nxR.size
    d = np.zeros((nt, NumRec), np.float32, order="F")
    F = np.zeros((ny, nx, nt), np.float32, order="F")
    S.solver_ac_forward_ram(0, 1, NumRec, nxS, nyS, nxR, nyR, src, m.dx, dt, ny, nx, ctrue, 1, nt, nt, F, d,
                            par=par, host_fields=False, ctx=ctx)
    same_gather = (abs(float(np.sum(d.astype(np.float64) ** 2)) - float(fx["data_sumsq"])) <=
                   1e-12 * float(fx["data_sumsq"]) and
                   np.array_equal(d.reshape(-1, order="F")[fx["data_pick"]], fx["data_sample"]))
    Sadj = np.asfortranarray(d[::-1, :])
    Rec = np.zeros((nt, NumRec), np.float32, order="F")
    k = np.zeros((ny, nx), np.float32, order="F")
    S.solver_ac_forward_ram(0, 1, NumRec, nxS, nyS, nxR, nyR, src, m.dx, dt, ny, nx, ctmp, 1, nt, nt, F, Rec,
                            par=par, host_fields=False, ctx=ctx)
    S.solver_ac_backward_ram(0, 1, NumRec, nxS, nyS, nxR, nyR, Sadj, m.dx, dt, ny, nx, ctmp, 1, nt, nt, F, k,
                             par=par, host_fields=False, ctx=ctx)
    return {"workload": f"single401 shot, {ny}x{nx} window, {nt} steps", "reference": "FP64 oracle fixture",
            "gather_matches_fixture": bool(same_gather), "image_rel_l2_vs_fp64": rel_l2(k, fx["image_fp64"]),
            "traces_rel_l2_vs_fp64": rel_l2(Rec, fx["traces_fp64"])}


ACC_BAR = 1e-4  # image rel-L2 against the FP64 oracle that the batched/served forms are held to


def accuracy_line(ctx, torch, dev, W, geos, d_mig, d_src, obs, cfg):
    """Shot 0 of the line at its full record length (601 x 715 window, nt 6000): the gather
    modelled on the device (true minus water-only model), migrated by the plain form and by
    the engine (launch shape `cfg`, one resident shot), and the images and the plain form's
    traces compared with the FP64 oracle's (tests/golden/accuracy_mcs2001_fp64.npz)."""
    import numpy as np

    from psdm_amd.engine import Engine
    from psdm_amd.solvers import backward_device, forward_device
    from tests.cases import rel_l2

    fx = np.load(os.path.join(ROOT, "tests", "golden", "accuracy_mcs2001_fp64.npz"))
    m, sh, nt = W["model"], W["shots"][0], W["nt"]
    assert int(fx["nt"]) == nt and int(fx["x_i"]) == sh.x_i and int(fx["nx_w"]) == sh.nx_w
    ny = m.nz
    d = obs[0].cpu().numpy()  # (NumRec, nt): Data(nt, NumRec) in Fortran order
    sumsq = float(np.sum(d.astype(np.float64) ** 2))
    same_gather = (abs(sumsq - float(fx["data_sumsq"])) <= 1e-12 * float(fx["data_sumsq"]) and
                   np.array_equal(d.reshape(-1)[fx["data_pick"]], fx["data_sample"]))
    win = d_mig[sh.x_i - 1:sh.x_f]
    d_rec = torch.empty((sh.nxRec.size, nt), device=dev)
    d_img = torch.zeros((sh.nx_w, ny), device=dev)
    forward_device(ctx, geos[0], win, d_src, d_rec, key=-7)
    backward_device(ctx, geos[0], win, obs[0].flip(1).contiguous(), d_img, key=-7)
    ctx.sync()
    img_plain = d_img.cpu().numpy().T
    rec = d_rec.cpu().numpy().T[:, fx["trace_channels"]]
    d_img2 = torch.zeros((sh.nx_w, ny), device=dev)
    eng = Engine(ctx, d_mig, ny, m.nx, nslots=1, max_nx=sh.nx_w, max_nt=nt, dPML=W["dPML"], fs=W["fs"], dmodel=m.dx,
                 dt=W["dt"], tile_x=cfg["tile_x"], stream_w=cfg.get("stream_w", 0), groups=1,
                 chunk_steps=cfg.get("chunk_steps", 64), **({"steps_per_launch": cfg["steps_per_launch"]}
                                                          if cfg.get("steps_per_launch", 1) != 1 else {}))
    eng.submit(**job_of(W, 0, d_src, obs), d_image_out=d_img2)
    eng.drain()
    eng.close()
    ctx.sync()
    img_batched = d_img2.cpu().numpy().T
    e_plain, e_batched = rel_l2(img_plain, fx["image_fp64"]), rel_l2(img_batched, fx["image_fp64"])
    # the opt-in zero-sum weights (psdm_rtm_set_weights): the same shot, plain and batched
    ctx.set_weights("zero_sum")
    try:
        d_rec_z = torch.empty((sh.nxRec.size, nt), device=dev)
        d_img_z = torch.zeros((sh.nx_w, ny), device=dev)
        forward_device(ctx, geos[0], win, d_src, d_rec_z, key=-7)
        backward_device(ctx, geos[0], win, obs[0].flip(1).contiguous(), d_img_z, key=-7)
        ctx.sync()
    finally:
        ctx.set_weights("reference")
    d_img_zb = torch.zeros((sh.nx_w, ny), device=dev)
    eng = Engine(ctx, d_mig, ny, m.nx, nslots=1, max_nx=sh.nx_w, max_nt=nt, dPML=W["dPML"], fs=W["fs"], dmodel=m.dx,
                 dt=W["dt"], tile_x=cfg["tile_x"], stream_w=cfg.get("stream_w", 0), groups=1,
                 chunk_steps=cfg.get("chunk_steps", 64), weights="zero_sum")
    eng.submit(**job_of(W, 0, d_src, obs), d_image_out=d_img_zb)
    eng.drain()
    eng.close()
    ctx.sync()
    img_z, img_zb = d_img_z.cpu().numpy().T, d_img_zb.cpu().numpy().T
    e_z, e_zb = rel_l2(img_z, fx["image_fp64"]), rel_l2(img_zb, fx["image_fp64"])
    zero_sum = {"plain": {"image_rel_l2_vs_fp64": e_z,
                          "traces_rel_l2_vs_fp64": rel_l2(d_rec_z.cpu().numpy().T[:, fx["trace_channels"]],
                                                          fx["traces_fp64"])},
                "batched": {"image_rel_l2_vs_fp64": e_zb},
                "batched_image_equals_plain": bool(np.array_equal(img_z, img_zb)),
                "image_rel_l2_vs_reference_form": rel_l2(img_z, img_plain),
                "within_bar": bool(max(e_z, e_zb) <= ACC_BAR),
                "note": "opt-in zero-sum second-derivative weights: the FP32 weights of the reference form sum to "
                        "-5.9e-8 instead of 0 (a spurious mass term whose phase error grows with the record); the "
                        "zero-sum form removes it, and no longer reproduces the reference's FP32 output bit for bit"}
    return {"workload": f"mcs2001 shot 0, {ny}x{sh.nx_w} window, nt {nt} (full record)",
            "reference": "FP64 oracle fixture (tests/golden/accuracy_mcs2001_fp64.npz)",
            "gather_matches_fixture": bool(same_gather),
            "plain": {"image_rel_l2_vs_fp64": e_plain, "traces_rel_l2_vs_fp64": rel_l2(rec, fx["traces_fp64"])},
            "batched": {"image_rel_l2_vs_fp64": e_batched},
            "batched_image_equals_plain": bool(np.array_equal(img_plain, img_batched)),
            "bar": ACC_BAR, "within_bar": bool(max(e_plain, e_batched) <= ACC_BAR),
            "note": "both forms reproduce the reference's FP32 arithmetic bit for bit; their distance to FP64 is set by "
                    "the FP32 rounding of the reference's FD weights (an FP64 run with FP32-rounded weights is 1.3e-4 "
                    "away), not by rounding that accumulates in the time loop",
            "zero_sum": zero_sum}


def model_gathers(ctx, torch, dev, W, geos_model, d_true, d_wat, d_src, nshots):
    """Observed gathers: modelled on the true and on the water-only model and subtracted
    (direct wave removed); natural time order, (NumRec, nt) = Data(nt, NumRec)."""
    from psdm_amd.solvers import forward_device

    shots, nt = W["shots"], W["nt"]
    nrec = shots[0].nxRec.size
    d_a = torch.empty((nrec, nt), device=dev)
    d_b = torch.empty((nrec, nt), device=dev)
    obs = {}
    for k in range(nshots):
        sh = shots[k]
        forward_device(ctx, geos_model[k], d_true[sh.x_i - 1:sh.x_f], d_src, d_a, key=-1)
        forward_device(ctx, geos_model[k], d_wat[sh.x_i - 1:sh.x_f], d_src, d_b, key=-1)
        ctx.sync()
        obs[k] = (d_a - d_b).contiguous()
        torch.cuda.synchronize()
    return obs


def set_plain_shape(ctx, cfg):
    """Launch shape of the one-shot solver from {"tile_x": .., "stream_w": ..}."""
    ctx.set_launch(tile_x=int(cfg.get("tile_x", 8)))
    ctx.set_stream_width(int(cfg.get("stream_w", 2)))
    return {"tile_x": int(cfg.get("tile_x", 8)), "stream_w": int(cfg.get("stream_w", 2)), "given": True}


def run_plain(ctx, torch, dev, W, geos, d_mig, d_src, obs, steps, warmup, timing_every, tune=True, shape=None):
    """One shot per call: forward (snapshots to HBM) + backward + FP64 stack per shot.
    With `shape` ({"tile_x", "stream_w"}) that launch shape is used; else with `tune` it is
    first chosen on a 300-step probe of shot 0's window."""
    from psdm_amd.solvers import DeviceShot, backward_device, forward_device

    m, shots, nt = W["model"], W["shots"], W["nt"]
    ny = m.nz
    nrec = shots[0].nxRec.size
    rev = {k: v.flip(1).contiguous() for k, v in obs.items()}  # BackSource = Data(nt:1:-1)
    d_kernel = torch.zeros((max(s.nx_w for s in shots), ny), device=dev)
    psdm_sum = torch.zeros((m.nx, ny), dtype=torch.float64, device=dev)
    fold = torch.zeros((m.nx, ny), dtype=torch.int32, device=dev)
    d_rec = torch.empty((nrec, nt), device=dev)
    torch.cuda.synchronize()
    tuned = None
    if shape:
        tuned = {"best": set_plain_shape(ctx, shape)}
    elif tune:
        from psdm_amd.autotune import tune_plain

        npr, sh = 300, shots[0]
        pgeo = DeviceShot(ny, sh.nx_w, npr, W["dPML"], W["fs"], m.dx, W["dt"], [sh.nxSou - sh.x_i + 1], [sh.nySou],
                          sh.nxRec - sh.x_i + 1, sh.nyRec)
        win = d_mig[sh.x_i - 1:sh.x_f]
        p_src = d_src[:, :npr].contiguous()
        p_rec = torch.empty((nrec, npr), device=dev)
        p_obs = rev[0][:, :npr].contiguous()
        torch.cuda.synchronize()

        def probe():
            forward_device(ctx, pgeo, win, p_src, p_rec, key=-22)
            backward_device(ctx, pgeo, win, p_obs, d_kernel, key=-22)

        best, trials = tune_plain(ctx, probe, log=log)
        d_kernel.zero_()
        torch.cuda.synchronize()
        tuned = {"best": best, "trials": trials}

    def migrate(k):
        sh = shots[k % len(shots)]
        win = d_mig[sh.x_i - 1:sh.x_f]
        forward_device(ctx, geos[k % len(shots)], win, d_src, d_rec, key=k)
        backward_device(ctx, geos[k % len(shots)], win, rev[k % len(shots)], d_kernel, key=k)
        ctx.check(ctx._L.psdm_rtm_stack_image(ctx.h, d_kernel.data_ptr(), ny, sh.nx_w, sh.x_i,
                                              psdm_sum.data_ptr(), fold.data_ptr(), m.nx), "stack_image")

    log(f"[bench] plain: warmup {warmup} shot(s)")
    for k in range(warmup):
        migrate(k)
    ctx.sync()
    psdm_sum.zero_()
    fold.zero_()
    torch.cuda.synchronize()
    ctx.set_timing(timing_every)
    log(f"[bench] plain: timing {steps} shots")
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        migrate(k)
    ctx.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n_ev, mean_ms = ctx.step_stats()
    ctx.set_timing(0)
    return dict(elapsed=elapsed, n_ev=n_ev, mean_ms=mean_ms, shots=[(k % len(shots)) for k in range(warmup, warmup + steps)],
                psdm_sum=psdm_sum, fold=fold, launch=tuned)


def run_plain_wide(ctx, torch, dev, nt, timing_every, tune=True, shape=None):
    """The plain form on the widest window of the served stream: one WAS gather (801 x 5102
    window of the 801 x 6001 @ 10 m model, 1000 receivers, dt 1 ms, FP32 snapshots every
    step) over its first nt steps, forward + backward.  One step of this window is large
    enough to fill the chip; returns the step kernel's roofline."""
    import numpy as np

    from psdm_amd.solvers import DeviceShot, backward_device, forward_device, ricker_was
    from psdm_amd.workloads import step_bytes, updates_per_shot, was_gathers

    G = was_gathers(1, seed=2)
    m, sh = G["model"], G["shots"][0]
    ny, nxw, dt, dPML = m.nz, sh.nx_w, G["dt"], G["dPML"]
    geo = DeviceShot(ny, nxw, nt, dPML, 0, m.dx, dt, [sh.nxSou - sh.x_i + 1], [sh.nySou], sh.nxRec - sh.x_i + 1,
                     sh.nyRec)
    d_mig = torch.from_numpy(np.ascontiguousarray(m.migration[:, sh.x_i - 1:sh.x_f].T)).to(dev)
    d_src = torch.from_numpy(ricker_was(nt, dt, G["f0"]).reshape(1, nt)).to(dev)
    g = torch.Generator(device=dev)
    g.manual_seed(9)
    d_obs = 1e-3 * torch.randn((sh.nxRec.size, nt), generator=g, device=dev)
    d_rec = torch.empty((sh.nxRec.size, nt), device=dev)
    d_img = torch.zeros((nxw, ny), device=dev)
    torch.cuda.synchronize()
    tuned = None
    if shape:
        tuned = {"best": set_plain_shape(ctx, shape)}
    elif tune:
        from psdm_amd.autotune import tune_plain

        npr = 200
        pgeo = DeviceShot(ny, nxw, npr, dPML, 0, m.dx, dt, [sh.nxSou - sh.x_i + 1], [sh.nySou], sh.nxRec - sh.x_i + 1,
                          sh.nyRec)
        p_src = d_src[:, :npr].contiguous()
        p_rec = torch.empty((sh.nxRec.size, npr), device=dev)
        p_obs = d_obs[:, :npr].contiguous()
        torch.cuda.synchronize()

        def probe():
            forward_device(ctx, pgeo, d_mig, p_src, p_rec, key=-12)
            backward_device(ctx, pgeo, d_mig, p_obs, d_img, key=-12)

        best, trials = tune_plain(ctx, probe, log=log)
        d_img.zero_()
        tuned = {"best": best, "trials": trials}
    forward_device(ctx, geo, d_mig, d_src, d_rec, key=-11)  # warm-up (allocations)
    backward_device(ctx, geo, d_mig, d_obs, d_img, key=-11)
    ctx.sync()
    ctx.set_timing(timing_every)
    t0 = time.perf_counter()
    forward_device(ctx, geo, d_mig, d_src, d_rec, key=-11)
    backward_device(ctx, geo, d_mig, d_obs, d_img, key=-11)
    ctx.sync()
    el = time.perf_counter() - t0
    n_ev, mean_ms = ctx.step_stats()
    ctx.set_timing(0)
    b = 0.5 * (step_bytes(ny, nxw, dPML) + step_bytes(ny, nxw, dPML, backward=True))
    ach = b / (mean_ms * 1e-3) / 1e9
    return {"workload": f"WAS gather, {ny}x{nxw} window, first {nt} of 12000 steps (dt 1 ms), FP32 snapshots",
            "launch": tuned,
            "gpt_updates_per_s": round(updates_per_shot(ny, nxw, nt, dPML) / el / 1e9, 3),
            "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4), "kernel": "rtm_step_kernel",
                         "mean_kernel_ms": round(mean_ms, 5), "sampled_launches": n_ev,
                         "algorithmic_bytes_per_launch": round(b)}}


def run_batched(ctx, torch, dev, W, d_mig, d_src, obs, steps, warmup, cfg):
    """Many shots per launch: the engine keeps `nslots` shots resident."""
    from psdm_amd.engine import Engine

    m, shots, nt = W["model"], W["shots"], W["nt"]
    ny = m.nz
    psdm_sum = torch.zeros((m.nx, ny), dtype=torch.float64, device=dev)
    fold = torch.zeros((m.nx, ny), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    eng = Engine(ctx, d_mig, ny, m.nx, nslots=cfg["nslots"], max_nx=max(s.nx_w for s in shots), max_nt=nt,
                 dPML=W["dPML"], fs=W["fs"], dmodel=m.dx, dt=W["dt"], tile_x=cfg["tile_x"], grid=cfg["grid"],
                 stream_w=cfg.get("stream_w", 0), groups=cfg.get("groups", 0),
                 snap_format=cfg.get("snap_format", "f32"),
                 store_snap=cfg.get("store_snap", 1), kahan=cfg.get("kahan", False),
                 chunk_steps=cfg.get("chunk_steps", 64), d_psdm_sum=psdm_sum, d_cov_sum=fold,
                 steps_per_launch=cfg.get("steps_per_launch", 1), weights=cfg.get("weights", "reference"))
    for k in range(warmup, warmup + steps):  # the shots the plain form times
        eng.submit(**job_of(W, k % len(shots), d_src, obs))
    ctx.sync()
    torch.cuda.synchronize()
    log(f"[bench] batched: {steps} shots, {cfg}")
    t0 = time.perf_counter()
    eng.drain()
    ctx.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st = eng.stats()
    launch_ms, n_launch = eng.timing()
    eng.close()
    return dict(elapsed=elapsed, stats=st, launch_ms=launch_ms, n_launch=n_launch, psdm_sum=psdm_sum, fold=fold)


def job_of(W, k, d_src, obs):
    sh = W["shots"][k]
    return dict(x_i=sh.x_i, nx_w=sh.nx_w, nt=W["nt"], nxSou=sh.nxSou - sh.x_i + 1, nySou=sh.nySou,
                nxRec=sh.nxRec - sh.x_i + 1, nyRec=sh.nyRec, d_wavelet=d_src[0], d_data=obs[k])


# ---------------------------------------------------------------------------------
# Other workloads (not the default line): uneven 256-shot line (batched vs plain),
# mixed-precision snapshots on it, and the served stream.  Each prints its own JSON line.


def interior_ld(ny):
    return -(-ny // 64) * 64


def snap_store_bytes(ny, nx_w, nt, store_snap, esize, ntiles=0):
    """Device bytes of one shot's snapshot store (rows padded to a multiple of 64, plus
    one scale per snapshot tile for 16-bit formats)."""
    nsnap = (nt - 1) // store_snap
    b = (nsnap * interior_ld(ny) * nx_w * esize + 255) // 256 * 256
    return b + (nsnap * ntiles * 4 if esize == 2 else 0)


def random_gathers(torch, dev, shots, nts, seed):
    """Seeded noise gathers (NumRec, nt) on the device: throughput does not depend on the
    recorded values."""
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    return {k: 1e-3 * torch.randn((sh.nxRec.size, nts[k]), generator=g, device=dev) for k, sh in enumerate(shots)}


def uneven_jobs(W, d_src, obs, idx, nt_override=None):
    out = []
    for k in idx:
        sh = W["shots"][k]
        nt = nt_override or W["nt"][k]
        out.append(dict(x_i=sh.x_i, nx_w=sh.nx_w, nt=nt, nxSou=sh.nxSou - sh.x_i + 1, nySou=sh.nySou,
                        nxRec=sh.nxRec - sh.x_i + 1, nyRec=sh.nyRec, d_wavelet=d_src[0], d_data=obs[k]))
    return out


def run_engine_jobs(ctx, torch, dev, d_mig, ny, nx_prop, jobs, W, cfg, budget, images=None):
    """Migrates `jobs` through one engine; returns elapsed seconds and stats.  `images`:
    optional list of device tensors (nx_w, ny) receiving each job's image."""
    from psdm_amd.engine import Engine

    eng = Engine(ctx, d_mig, ny, nx_prop, nslots=cfg["nslots"], max_nx=max(j["nx_w"] for j in jobs),
                 max_nt=max(j["nt"] for j in jobs), dPML=W["dPML"], fs=W["fs"], dmodel=W["model"].dx, dt=W["dt"],
                 tile_x=cfg["tile_x"], stream_w=cfg.get("stream_w", 0), grid=cfg.get("grid", 0),
                 groups=cfg.get("groups", 0),
                 snap_format=cfg.get("snap_format", "f32"),
                 store_snap=cfg.get("store_snap", 1), kahan=cfg.get("kahan", False),
                 chunk_steps=cfg.get("chunk_steps", 64), hbm_budget=budget, weights=cfg.get("weights", "reference"),
                 steps_per_launch=cfg.get("steps_per_launch", 1))
    for q, j in enumerate(jobs):
        eng.submit(**j, d_image_out=images[q] if images is not None else None)
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.drain()
    ctx.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    st = eng.stats()
    eng.close()
    return el, st


def bench_uneven(args, mixed=False):
    """Uneven 256-shot line (windows 600-2000 columns, nt 4000/6000/8000): the batched
    engine vs the plain one-shot form on the same shots; with `mixed`, FP32 / FP16 / BF16
    snapshots every 4th step (Kahan image) against FP32 every step: throughput, snapshot
    HBM and speedup (their accuracy against the FP64 oracle is measured by
    tests/accuracy_report.py)."""
    import numpy as np
    import torch

    from psdm_amd._native import Context
    from psdm_amd.solvers import DeviceShot, backward_device, forward_device, ricker_mcs
    from psdm_amd.workloads import step_bytes, uneven_line, updates_per_shot

    W = uneven_line(256, seed=0)
    m, shots = W["model"], W["shots"][:args.shots]
    nts = W["nt"][:args.shots]
    ny, dPML = m.nz, W["dPML"]
    dev = torch.device("cuda:0")
    ctx = Context(0)
    d_mig = torch.from_numpy(np.ascontiguousarray(m.migration.T)).to(dev)
    ntmax = max(nts)
    d_src = torch.from_numpy(ricker_mcs(ntmax, W["dt"], W["f0"], 1.2 / W["f0"]).reshape(1, ntmax)).to(dev)
    obs = random_gathers(torch, dev, shots, nts, seed=0)
    torch.cuda.synchronize()
    jobs = uneven_jobs(W, d_src, obs, range(len(shots)))
    upd = sum(updates_per_shot(ny, sh.nx_w, nts[k], dPML) for k, sh in enumerate(shots))
    budget = int(args.hbm_gb * 2 ** 30)
    log(f"[bench] uneven line: {len(shots)} shots, windows {min(s.nx_w for s in shots)}-"
        f"{max(s.nx_w for s in shots)}, {upd / 1e12:.2f} T updates, snapshot budget {args.hbm_gb} GB")

    cfg = dict(nslots=args.slots, tile_x=16, stream_w=4, grid=0, groups=2, chunk_steps=64)
    tuned = None
    if not args.no_tune:
        # launch shape and steps per launch (temporal blocking) chosen by the autotuner on the
        # line's own windows, at the given slot count and snapshot budget
        from psdm_amd.autotune import tune

        best, trials = tune(ctx, d_mig, ny, m.nx, jobs, max(s.nx_w for s in shots), max(nts), dPML, W["fs"], m.dx,
                            W["dt"], slots=[args.slots], steps_per_launch=(1, 2), hbm_budget=budget, log=log)
        cfg.update(tile_x=best["tile_x"], stream_w=best["stream_w"], grid=best["grid"], groups=best["groups"],
                   steps_per_launch=best["steps_per_launch"])
        tuned = {"best": best, "trials": trials}
    # mixed: the 16-bit forms, and FP32 / FP16 every 4th step with the zero-sum weights (the
    # forms that stay within 1e-4 of the FP64 oracle at nt 8000, tests/accuracy_report.py)
    forms = [("f32", 1, False, "reference")] + ([("f32", 4, True, "reference"), ("f16", 4, True, "reference"),
                                                 ("bf16", 4, True, "reference"), ("f32", 4, True, "zero_sum"),
                                                 ("f16", 4, True, "zero_sum")] if mixed else [])
    results = {}
    for fmt, ss, kahan, weights in forms:
        c = dict(cfg, snap_format=fmt, store_snap=ss, kahan=kahan, weights=weights)
        if fmt != "f32" or weights != "reference":
            c["steps_per_launch"] = 1  # two steps per launch take FP32 snapshots and the reference weights only
        log(f"[bench] batched {fmt} every {ss}, {weights} weights: {c}")
        el, st = run_engine_jobs(ctx, torch, dev, d_mig, ny, m.nx, jobs, W, c, budget)
        esize = 4 if fmt == "f32" else 2
        sb = sum(snap_store_bytes(ny, sh.nx_w, nts[k], ss, esize) for k, sh in enumerate(shots))
        byts = sum(0.5 * (step_bytes(ny, sh.nx_w, dPML, esize, ss) + step_bytes(ny, sh.nx_w, dPML, esize, ss, True))
                   * 2 * (nts[k] - 1) for k, sh in enumerate(shots))
        r = {"snap_format": fmt, "store_snap": ss, "kahan": kahan, "weights": weights,
             "shots_per_s": round(len(shots) / el, 4),
             "gpt_updates_per_s": round(upd / el / 1e9, 3), "elapsed_s": round(el, 3),
             "snapshot_bytes_total": int(sb), "hbm_roofline_frac": round(byts / el / (HBM_PEAK_GBS * 1e9), 4),
             "steps_issued": st["steps_issued"], "slot_steps": st["slot_steps"]}
        results[f"{fmt}_ss{ss}" + ("" if weights == "reference" else "_zero_sum")] = r
        log(f"[bench]   {r['shots_per_s']} shots/s, {r['gpt_updates_per_s']} Gupd/s, snapshots "
            f"{sb / 2 ** 30:.1f} GiB")

    plain = None
    if not args.no_plain:
        log(f"[bench] plain: {len(shots)} shots one by one")
        ctx.set_launch(tile_x=args.tile_x)
        d_kernel = torch.zeros((max(s.nx_w for s in shots), ny), device=dev)
        d_rec = torch.empty((max(s.nxRec.size for s in shots), ntmax), device=dev)
        geos = [DeviceShot(ny, sh.nx_w, nts[k], dPML, W["fs"], m.dx, W["dt"], [sh.nxSou - sh.x_i + 1], [sh.nySou],
                           sh.nxRec - sh.x_i + 1, sh.nyRec) for k, sh in enumerate(shots)]
        psdm_sum = torch.zeros((m.nx, ny), dtype=torch.float64, device=dev)
        fold = torch.zeros((m.nx, ny), dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k, sh in enumerate(shots):
            win = d_mig[sh.x_i - 1:sh.x_f]
            rev_k = obs[k].flip(1).contiguous()  # BackSource = Data(nt:1:-1)
            forward_device(ctx, geos[k], win, d_src, d_rec, key=k)
            backward_device(ctx, geos[k], win, rev_k, d_kernel, key=k)
            ctx.check(ctx._L.psdm_rtm_stack_image(ctx.h, d_kernel.data_ptr(), ny, sh.nx_w, sh.x_i,
                                                  psdm_sum.data_ptr(), fold.data_ptr(), m.nx), "stack_image")
        ctx.sync()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        plain = {"shots_per_s": round(len(shots) / el, 4), "gpt_updates_per_s": round(upd / el / 1e9, 3),
                 "elapsed_s": round(el, 3), "tile_x": args.tile_x}
        log(f"[bench] plain {plain['shots_per_s']} shots/s")
        if not args.no_wide:
            plain["wide_window"] = run_plain_wide(ctx, torch, dev, args.wide_steps, args.timing_every)
            log(f"[bench] plain, wide window: roofline {plain['wide_window']['roofline']['frac']}")

    base = results["f32_ss1"]
    if mixed:
        for key, r in results.items():
            r["speedup_vs_f32_every_step"] = round(r["shots_per_s"] / base["shots_per_s"], 3)
            r["snapshot_hbm_saving_vs_f32_every_step"] = round(base["snapshot_bytes_total"] /
                                                               r["snapshot_bytes_total"], 3)
    head = results["f16_ss4"] if mixed else base
    out = {
        "metric": "migrated shots/s + Gpt-updates/s, batched vs plain, at rel-L2<=1e-4; p99 latency",
        "value": head["shots_per_s"], "unit": "shots/s", "n_gpus": 1, "steps": len(shots), "warmup": 0,
        "ms_per_step": round(1e3 / head["shots_per_s"], 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded 4001x601 model; seeded noise gathers)",
        "config": {"workload": f"uneven line: {len(shots)} shots on a 40 km 601-deep model @10 m, windows "
                               f"600-2000 columns, nt in {{4000,6000,8000}}, dt 0.8 ms, dPML 40",
                   "form": "batched" + (" FP16 snapshots every 4th step" if mixed else ""),
                   "slots": args.slots, "launch": {k: cfg[k] for k in ("tile_x", "stream_w", "grid", "groups")},
                   "steps_per_launch": cfg.get("steps_per_launch", 1), "snapshot_budget_gb": args.hbm_gb,
                   "autotune": tuned},
        "gpt_updates_per_s": head["gpt_updates_per_s"],
        "forms": results, "plain": plain,
        "speedup_batched_vs_plain": round(base["shots_per_s"] / plain["shots_per_s"], 3) if plain else None,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_condition(args):
    """Gather conditioning on the device (psdm_rtm_condition): `--steps` MCS gathers of 480
    traces x 2401 samples at 2 ms (4.8 s records), band-pass 5-40 Hz, Akima resampling to
    the 0.8 ms propagation step (6000 samples), cut before the water-bottom reflection and
    a 0.2 s half-cosine taper (calculate_PSDM_RTM.f90:229-358).  Inputs resident in HBM;
    one JSON line.  The CPU baseline runs the reference's own routines (oracle/_ref)."""
    import numpy as np

    from psdm_amd.conditioning import cut_indices, taper_lengths

    nt_data, dt_data, nrec = 2401, 0.002, 480
    dt_prop, nt_prop2 = 0.0008, 6000
    typef, f1, f2 = 3, 5.0, 40.0
    N_taper, _ = taper_lengths(0.2, 0.0, dt_prop)
    cpu = None
    if not args.no_cpu:  # before anything touches the GPU (the pool spawns processes)
        from oracle import cpu_baseline

        per_rank = 96
        log(f"[bench] condition cpu baseline: {args.cpu_procs} procs x {per_rank} traces")
        r = cpu_baseline.run_conditioning(nt_data, dt_data, per_rank, nt_prop2, dt_prop, typef, f1, f2, N_taper,
                                          args.cpu_procs)
        cpu = {"value": round(r["traces_per_s"] / nrec, 4), "unit": "gathers/s", "cores": r["cores"],
               "kind": r["kind"],
               "sample": f"{r['cores']} processes, each conditioning {per_rank} traces of {nt_data} samples "
                         f"(freq_filter band-pass {f1:g}-{f2:g} Hz with FFTPACK over the x24 zero-padded trace, "
                         f"interpolation_akima {dt_data * 1e3:g} -> {dt_prop * 1e3:g} ms, taper_pad_trace); "
                         f"gathers/s = traces/s / {nrec} (slowest process {r['rank_s']:.2f} s)"}
        log(f"[bench] condition cpu baseline {cpu['value']} gathers/s")

    import torch

    from psdm_amd import conditioning as CD
    from psdm_amd._native import Context
    from psdm_amd.workloads import line_workload

    dev = torch.device("cuda:0")
    ctx = Context(0)
    W = line_workload("mcs2001")
    m, shots = W["model"], W["shots"]
    bat = np.full(m.nx, m.water_depth, np.float32)
    n_src = 16  # distinct device gathers cycled through the timed region
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    raw = torch.randn((n_src, nrec, nt_data), generator=gen, device=dev)
    outs = torch.empty((n_src, nrec, nt_prop2), device=dev)
    sfs = [cut_indices(sh.nxSou, sh.nySou, sh.nxRec, sh.nyRec, bat, m.dx, dt_prop, "sf")
           for sh in shots[:n_src]]
    kw = dict(dt_data=dt_data, nt_prop2=nt_prop2, dt_prop=dt_prop, typef=typef, f_init=f1, f_final=f2,
              taper_option=1, N_taper=N_taper)
    for k in range(max(args.warmup, 1)):  # the first call designs the filter taps
        CD.condition(ctx, raw[k % n_src], nt_sf=sfs[k % n_src], out=outs[k % n_src], **kw)
    ctx.set_timing(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        CD.condition(ctx, raw[k % n_src], nt_sf=sfs[k % n_src], out=outs[k % n_src], **kw)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    f_ms, r_ms, n_timed = CD.timing(ctx)
    ctx.set_timing(0)
    flops = 2.0 * nt_data * nt_data * nrec
    ach = flops / (f_ms * 1e-3) / 1e12
    rbytes = 4.0 * nrec * (nt_data + nt_prop2)
    acc = None if args.no_accuracy else condition_accuracy(ctx)
    tpath = os.path.join(ROOT, "profiles", "traffic_cond_toeplitz_kernel_mcs.md")
    traffic = None
    if os.path.exists(tpath):
        import re

        mt = re.search(r"^traffic_bytes_per_launch = (\d+)", open(tpath).read(), re.M)
        traffic = int(mt.group(1)) if mt else None
    out = {
        "metric": "conditioned gathers/s (filter + resample + cut/taper)",
        "value": round(args.steps / elapsed, 3), "unit": "gathers/s", "n_gpus": 1, "steps": args.steps,
        "warmup": max(args.warmup, 1), "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Gaussian traces; cut_sf from the mcs2001 line's geometry and water depth)",
        "config": {"workload": f"MCS gathers: {nrec} traces x {nt_data} samples @ {dt_data * 1e3:g} ms -> "
                               f"{nt_prop2} @ {dt_prop * 1e3:g} ms, band-pass {f1:g}-{f2:g} Hz, "
                               f"cut_sf, taper {N_taper} samples"},
        "roofline": {"bound": "mfma", "achieved": round(ach, 2), "peak": 157.3, "unit": "TFLOP/s",
                     "frac": round(ach / 157.3, 4), "traffic": traffic, "kernel": "cond_toeplitz_kernel",
                     "mean_kernel_ms": round(f_ms, 5), "timed_launches": n_timed,
                     "algorithmic_flops_per_launch": flops,
                     "note": "Toeplitz product on the matrix cores (v_mfma_f32_32x32x2f32), 2*nt^2*traces "
                             "FLOP per gather; peak = MI355X dense FP32 157.3 TFLOP/s"},
        "resample_kernel": {"mean_kernel_ms": round(r_ms, 5), "algorithmic_bytes": rbytes,
                            "achieved_gbs": round(rbytes / (r_ms * 1e-3) / 1e9, 1) if r_ms else None},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(args.steps / elapsed / cpu["value"], 1) if cpu else None,
        "accuracy": acc,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def condition_accuracy(ctx):
    """Device conditioning of the golden cases (written from the reference's own routines,
    tests/golden/conditioning_ref.npz): rel-L2 where the filter runs, bit-equality else."""
    import numpy as np
    import torch

    from psdm_amd.conditioning import condition

    z = np.load(os.path.join(ROOT, "tests", "golden", "conditioning_ref.npz"))
    res = {}
    for name in [k[:-4] for k in z.files if k.endswith("_raw")]:
        p = dict(zip(z["param_names"], z[name + "_params"]))
        d = torch.from_numpy(np.ascontiguousarray(z[name + "_raw"].T)).to("cuda:0")
        o = condition(ctx, d, dt_data=float(p["dt_data"]), typef=int(p["typef"]), f_init=float(p["f_init"]),
                      f_final=float(p["f_final"]), nt_prop2=int(p["nt_prop2"]), dt_prop=float(p["dt_prop"]),
                      nt_taper=int(p["nt_taper"]), N_taper=int(p["N_taper"]), N_zeros=int(p["N_zeros"]),
                      taper_option=int(p["taper_option"]),
                      nt_sf=z[name + "_nt_sf"] if name + "_nt_sf" in z.files else None,
                      nt_cut=z[name + "_nt_cut"] if name + "_nt_cut" in z.files else None)
        o = o.cpu().numpy().T
        ref = z[name + "_out"]
        res[name] = {"rel_l2_vs_reference": float(np.linalg.norm(o - ref) / np.linalg.norm(ref)),
                     "bit_equal": bool(np.array_equal(o, ref))}
    return res


def bench_poststack(args):
    """The steps after the stack on the device: normalize (PSDM = PSDM_sum / Cov_sum,
    auxiliar_PSDM.f90:1-20) and the depth-to-two-way-time conversion of the image and the
    velocity (psdm2twt, :23-91) for the 2001 x 601 @ 10 m line (TWT step 4 ms).  `--steps`
    conversions timed; one JSON line.  CPU baseline: the reference's linear_interpolation_1D
    (oracle/_ref) over a share of the columns per process."""
    import numpy as np

    nz, nx, dd, dt = 601, 2001, 10.0, 0.004
    cpu = None
    if not args.no_cpu:
        from oracle import cpu_baseline

        per_rank = 500
        r = cpu_baseline.run_twt(nz, per_rank, dd, dt, args.cpu_procs)
        cpu = {"value": round(r["cols_per_s"] / nx, 3), "unit": "images/s", "cores": r["cores"], "kind": r["kind"],
               "sample": f"{r['cores']} processes, each converting {per_rank} columns of {nz} depth samples "
                         f"(two-way time accumulation, then linear_interpolation_1D of velocity and image onto "
                         f"a {dt * 1e3:g} ms axis); images/s = columns/s / {nx} (slowest {r['rank_s']:.2f} s)"}
        log(f"[bench] poststack cpu baseline {cpu['value']} images/s")

    import torch

    from psdm_amd import poststack as PS
    from psdm_amd._native import Context
    from psdm_amd.workloads import line_workload

    dev = torch.device("cuda:0")
    ctx = Context(0)
    m = line_workload("mcs2001")["model"]
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    d_vel = torch.from_numpy(np.ascontiguousarray(m.migration.T)).to(dev)  # (nx, nz)
    d_sum = torch.randn((nx, nz), generator=gen, device=dev, dtype=torch.float64)
    d_fold = torch.randint(0, 40, (nx, nz), generator=gen, device=dev, dtype=torch.int32)
    d_img = torch.empty((nx, nz), device=dev)

    out = []

    def one():
        import ctypes as C

        ctx.check(ctx._L.psdm_rtm_normalize(ctx.h, C.c_void_p(d_sum.data_ptr()), C.c_void_p(d_fold.data_ptr()), nz,
                                            nx, C.c_void_p(d_img.data_ptr())), "psdm_rtm_normalize")
        return PS.psdm2twt(ctx, d_vel, d_img, dd, dt, out=out or None)

    # the first conversion sizes the outputs; the timed ones reuse them (one call each: the
    # times are computed anew every call, the length query is not repeated)
    vt, pt = one()
    nt = vt.shape[1]
    out.extend([torch.empty(nx * nt, device=dev), torch.empty(nx * nt, device=dev)])
    v2, p2 = one()
    assert torch.equal(v2, vt) and torch.equal(p2, pt)
    ctx.set_timing(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    a_ms, b_ms, n = PS.timing(ctx)
    ctx.set_timing(0)
    # dominant kernel: the time accumulation (reads the velocity, writes the times); the
    # interpolation reads the time, velocity and image columns and writes both outputs
    byts = 4.0 * (2 * nz * nx)
    ach = byts / (a_ms * 1e-3) / 1e9
    ibytes = 4.0 * (3 * nz * nx + 2 * nt * nx)
    out = {
        "metric": "post-stack images/s (normalize + depth-to-TWT of image and velocity)",
        "value": round(args.steps / elapsed, 2), "unit": "images/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (mcs2001 migration velocity; seeded FP64 stack and fold)",
        "config": {"workload": f"{nx} x {nz} image @ {dd:g} m -> {nt} TWT samples @ {dt * 1e3:g} ms"},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "twt_times_kernel",
                     "mean_kernel_ms": round(a_ms, 5), "timed_launches": n, "algorithmic_bytes_per_launch": byts,
                     "note": "each column's times are an FP32 running sum in depth order (the reference's order, "
                             "bit for bit): a serial chain of nz dependent additions per column, latency bound"},
        "interp_kernel": {"kernel": "twt_interp_kernel", "mean_kernel_ms": round(b_ms, 5),
                          "algorithmic_bytes": ibytes, "achieved_gbs": round(ibytes / (b_ms * 1e-3) / 1e9, 1)},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(args.steps / elapsed / cpu["value"], 1) if cpu else None,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_su(args):
    """The SU formats on either side of the migration: the raw records of the 64-shot line
    (480 traces x 2401 samples @ 2 ms per shot, big-endian SU, 302 MB) decoded from
    device-resident file bytes into the trace-major FP32 gathers in one call per step
    (get_shot_data, get_mcs_data.f90:150-214), and the final 2001 x 601 image encoded as an
    SU file (write_psdm, auxiliar.f90:275-299).  `--steps` line decodes timed; one JSON
    line.  CPU baseline: the reference's open_shot_mcs + get_shot_data (oracle/_ref) on a
    file in the page cache."""
    import numpy as np

    nshot, nrec, ns = 64, 480, 2401
    cpu = None
    if not args.no_cpu:
        from oracle import cpu_baseline

        reps = 150
        r = cpu_baseline.run_su(nrec, ns, reps, args.cpu_procs)
        cpu = {"value": round(r["gathers_per_s"], 2), "unit": "gathers/s", "cores": r["cores"], "kind": r["kind"],
               "sample": f"{r['cores']} processes, each reading one {nrec}-trace x {ns}-sample big-endian SU gather "
                         f"{reps} times through open_shot_mcs + get_shot_data (page cache; slowest "
                         f"{r['rank_s']:.2f} s)"}
        log(f"[bench] su cpu baseline {cpu['value']} gathers/s")

    import torch

    from psdm_amd import su
    from psdm_amd._native import Context

    dev = torch.device("cuda:0")
    ctx = Context(0)
    rng = np.random.default_rng(0)
    shot = b"".join(rng.integers(0, 1000, 60).astype(">i4").tobytes() + rng.standard_normal(ns).astype(">f4").tobytes()
                    for _ in range(nrec))
    host = np.frombuffer(shot * nshot, np.uint8)
    pinned = torch.empty(host.size, dtype=torch.uint8, pin_memory=True)
    pinned.numpy()[:] = host
    d_su = pinned.to(dev)
    ntr = nshot * nrec
    out = torch.empty((ntr, ns), device=dev)
    su.decode(ctx, d_su, ntr, ns, out=out)
    ok = bool(np.array_equal(out[:nrec].cpu().numpy(), np.frombuffer(shot, ">f4").reshape(nrec, 60 + ns)[:, 60:]))
    ctx.set_timing(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        su.decode(ctx, d_su, ntr, ns, out=out)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    dec_ms, _, ndec, _ = su.timing(ctx)
    # with the host-to-device copy of the file bytes from pinned memory (PCIe)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(3):
        d_su.copy_(pinned)
        su.decode(ctx, d_su, ntr, ns, out=out)
    torch.cuda.synchronize()
    h2d_s = (time.perf_counter() - t1) / 3
    # the same bytes moved by a plain device-to-device copy (torch's copy kernel): the
    # practical streaming rate the decode is compared with
    src = out.clone()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out.copy_(src)
    e0.record()
    for _ in range(20):
        out.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    copy_ms = e0.elapsed_time(e1) / 20
    del src
    # the final image as an SU file
    nz, nx = 601, 2001
    img = torch.randn((nx, nz), device=dev)
    enc = torch.empty(nx * (240 + 4 * nz), dtype=torch.uint8, device=dev)
    su.encode(ctx, img, 10.0, 10.0, out=enc)
    ctx.set_timing(0)
    ctx.set_timing(1)
    for _ in range(20):
        su.encode(ctx, img, 10.0, 10.0, out=enc)
    _, enc_ms, _, nenc = su.timing(ctx)
    ctx.set_timing(0)
    byts = 8.0 * ntr * ns  # read and write each sample once (headers are not read)
    ach = byts / (dec_ms * 1e-3) / 1e9
    tpath = os.path.join(ROOT, "profiles", "traffic_su_decode_kernel_line.md")
    traffic = None
    if os.path.exists(tpath):
        import re

        mt = re.search(r"^traffic_bytes_per_launch = (\d+)", open(tpath).read(), re.M)
        traffic = int(mt.group(1)) if mt else None
    ebytes = 4.0 * nx * nz + nx * (240 + 4 * nz)
    out_j = {
        "metric": "SU gathers decoded per second (device-resident file bytes -> FP32 gathers)",
        "value": round(nshot * args.steps / elapsed, 1), "unit": "gathers/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": round(1e3 * elapsed / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded big-endian SU records)",
        "config": {"workload": f"{nshot} shots x {nrec} traces x {ns} samples (2 ms, 4.8 s), big-endian SU, "
                               f"{host.size / 1e6:.1f} MB per line, one decode call per line"},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "su_decode_kernel",
                     "mean_kernel_ms": round(dec_ms, 5), "timed_launches": ndec, "algorithmic_bytes_per_launch": byts},
        "decode_matches_bytes": ok,
        "plain_copy": {"bytes": byts, "ms": round(copy_ms, 5), "achieved_gbs": round(byts / (copy_ms * 1e-3) / 1e9, 1),
                       "note": "torch device-to-device copy of the decoded gathers (same bytes read and written)"},
        "with_h2d": {"gathers_per_s": round(nshot / h2d_s, 1), "ms_per_line": round(1e3 * h2d_s, 3),
                     "note": "host-to-device copy of the line's bytes from pinned memory, then the decode"},
        "encode": {"kernel": "su_encode_kernel", "workload": f"{nx} x {nz} image -> SU", "mean_kernel_ms":
                   round(enc_ms, 5), "calls": nenc, "algorithmic_bytes": ebytes,
                   "achieved_gbs": round(ebytes / (enc_ms * 1e-3) / 1e9, 1)},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(nshot * args.steps / elapsed / cpu["value"], 1) if cpu else None,
    }
    print(json.dumps(out_j), flush=True)
    ctx.close()


def bench_kirchhoff(args):
    """Kirchhoff migration of the 64-shot line's shots (calculate_PSDM_kirchoff.f90:352-447):
    per shot the travel times in the 601 x 715 window from the source and the 480 receivers
    (481 fast-sweeping eikonal solves, one call) and the diffraction stack over 480 traces of
    6000 samples at 0.8 ms.  `--steps` shots timed; one JSON line.  CPU baseline: the
    oracle's C restatement (its eikonal is the reference's, bit for bit) on a sample."""
    import numpy as np

    from psdm_amd.workloads import line_workload

    W = line_workload("mcs2001")
    m, sh, dt, nt = W["model"], W["shots"][0], W["dt"], W["nt"]
    ny, nx, dd, nrec = m.nz, sh.nx_w, m.dx, sh.nxRec.size
    cpu = None
    if not args.no_cpu:
        from oracle import cpu_baseline

        r = cpu_baseline.run_kirchhoff(ny, nx, dd, nrec, nt, dt, 12, 12, args.cpu_procs)
        cpu = {"value": round(r["shots_per_s"], 5), "unit": "shots/s", "cores": r["cores"], "kind": r["kind"],
               "sample": f"{r['cores']} processes, each timing 12 eikonal solves and the imaging of 12 receivers in "
                         f"the {ny} x {nx} window; a shot = {nrec + 1} solves + {nrec} receivers "
                         f"({r['shot_s']:.1f} s per shot per process)"}
        log(f"[bench] kirchhoff cpu baseline {cpu['value']} shots/s")

    import torch

    from psdm_amd import kirchhoff as KI
    from psdm_amd._native import Context

    dev = torch.device("cuda:0")
    ctx = Context(0)
    d_vel = torch.from_numpy(np.ascontiguousarray(m.migration.T[sh.x_i - 1:sh.x_i - 1 + nx])).to(dev)  # (nx, ny)
    xs = np.concatenate([[sh.nxSou - sh.x_i + 1], sh.nxRec - sh.x_i + 1]).astype(np.int32)
    zs = np.concatenate([[sh.nySou], sh.nyRec]).astype(np.int32)
    tt = torch.empty((nrec + 1, nx, ny), device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    data = torch.randn((nrec, nt), generator=gen, device=dev)
    img = torch.empty((nx, ny), device=dev)

    def one():
        KI.travel_times(ctx, d_vel, dd, xs, zs, out=tt)
        KI.image(ctx, tt[0], tt[1:], data, nt, dd, dt, out=img)

    one()
    ctx.set_timing(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    eik_ms, img_ms, n_eik, n_img = KI.timing(ctx)
    ctx.set_timing(0)
    cells = (nrec + 1) * nx * ny
    # eikonal: per cell and sweep the old value and the velocity read, the new value written
    ebytes = 12.0 * 4 * cells
    ach = ebytes / (eik_ms * 1e-3) / 1e9
    ibytes = 4.0 * (nrec + 1) * nx * ny + 4.0 * nx * ny  # travel times read once, image written
    out = {
        "metric": "Kirchhoff-migrated shots/s (eikonal travel times + diffraction stack)",
        "value": round(args.steps / elapsed, 3), "unit": "shots/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded mcs2001 migration velocity; seeded traces)",
        "config": {"workload": f"mcs2001 shot: {ny} x {nx} window @ {dd:g} m, {nrec + 1} eikonal solves, "
                               f"{nrec} traces x {nt} samples @ {dt * 1e3:g} ms"},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "eikonal_fsm_kernel",
                     "mean_kernel_ms": round(eik_ms, 4), "timed_launches": n_eik, "algorithmic_bytes_per_launch": ebytes,
                     "note": "four sweeps of nx + ny - 1 dependent wavefront steps per solve, one barrier each: "
                             "bound by the step chain's latency, not by bandwidth"},
        "image_kernel": {"kernel": "kirchhoff_image_kernel", "mean_kernel_ms": round(img_ms, 4),
                         "pixel_receiver_pairs": nrec * nx * ny, "algorithmic_bytes": ibytes,
                         "achieved_gbs": round(ibytes / (img_ms * 1e-3) / 1e9, 1)},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(args.steps / elapsed / cpu["value"], 1) if cpu else None,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_served(args):
    """Served stream: Poisson arrivals at 4 rates; 70% MCS gathers (480 ch @12.5 m, 8 km
    aperture, 601-deep model @10 m, nt 6000, FP32 snapshots every step) and 30% WAS
    gathers (801 x 6001 model @10 m, 50 km offset, dt 1 ms, nt 12000, 8 Hz, FP16
    snapshots every 4th step); one engine per model, admission by each engine's HBM
    budget; per-rate p50/p99 latency (arrival -> image stacked), shots/s, occupancy."""
    import numpy as np
    import torch

    from psdm_amd._native import Context
    from psdm_amd.engine import Engine
    from psdm_amd.serve import Request, Server, poisson_arrivals
    from psdm_amd.solvers import ricker_mcs, ricker_was
    from psdm_amd.workloads import _mcs_shot, velocity_model, was_gathers

    dev = torch.device("cuda:0")
    ctx = Context(0)
    dx = 10.0
    mcs_model = velocity_model(601, 4001, dx, 1500.0, seed=1)
    added = (8000.0 - 150.0 - 479 * 12.5) / 2  # 8 km aperture
    rng = np.random.default_rng(5)
    n = args.requests
    nmcs = int(round(0.7 * n))
    mcs_shots = [_mcs_shot(q, float(rng.uniform(1000.0, 40000.0 - 8000.0)), dx, 4001, 480, 12.5, 150.0, 6.0, 8.0,
                           added) for q in range(nmcs)]
    WAS = was_gathers(n - nmcs, seed=2)
    was_model, was_shots = WAS["model"], WAS["shots"]
    nt_m, dt_m, nt_w, dt_w = 6000, 0.0008, WAS["nt"], WAS["dt"]
    d_vm = torch.from_numpy(np.ascontiguousarray(mcs_model.migration.T)).to(dev)
    d_vw = torch.from_numpy(np.ascontiguousarray(was_model.migration.T)).to(dev)
    s_m = torch.from_numpy(ricker_mcs(nt_m, dt_m, 15.0, 1.2 / 15.0)).to(dev)
    s_w = torch.from_numpy(ricker_was(nt_w, dt_w, 8.0)).to(dev)  # Ricker_was: t0 = 1.5/f0
    obs_m = random_gathers(torch, dev, mcs_shots, [nt_m] * nmcs, seed=3)
    obs_w = random_gathers(torch, dev, was_shots, [nt_w] * len(was_shots), seed=4)
    torch.cuda.synchronize()

    def mk_req(kind, q, sh, nt, d_s, d_o):
        job = dict(x_i=sh.x_i, nx_w=sh.nx_w, nt=nt, nxSou=sh.nxSou - sh.x_i + 1, nySou=sh.nySou,
                   nxRec=sh.nxRec - sh.x_i + 1, nyRec=sh.nyRec, d_wavelet=d_s, d_data=d_o)
        return kind, job

    pool_reqs = [mk_req("mcs", q, sh, nt_m, s_m, obs_m[q]) for q, sh in enumerate(mcs_shots)]
    pool_reqs += [mk_req("was", q, sh, nt_w, s_w, obs_w[q]) for q, sh in enumerate(was_shots)]
    order = np.random.default_rng(6).permutation(len(pool_reqs))
    gb = 2 ** 30
    mcs_cfg = dict(nslots=args.slots, max_nx=max(s.nx_w for s in mcs_shots), max_nt=nt_m, snap_format="f32",
                   store_snap=1, kahan=False, budget=int(args.hbm_gb * 0.62 * gb))
    was_cfg = dict(nslots=4, max_nx=max(s.nx_w for s in was_shots), max_nt=nt_w, snap_format="f16", store_snap=4,
                   kahan=True, budget=int(args.hbm_gb * 0.38 * gb))
    log(f"[bench] served: {n} requests ({nmcs} MCS windows ~{mcs_shots[0].nx_w}, {n - nmcs} WAS windows "
        f"{was_shots[0].nx_w}), rates {args.rates}")

    def engines():
        e = {"mcs": Engine(ctx, d_vm, 601, 4001, nslots=mcs_cfg["nslots"], max_nx=mcs_cfg["max_nx"],
                           max_nt=nt_m, dPML=40, fs=0, dmodel=dx, dt=dt_m, tile_x=16, snap_format="f32",
                           store_snap=1, hbm_budget=mcs_cfg["budget"]),
             "was": Engine(ctx, d_vw, 801, 6001, nslots=was_cfg["nslots"], max_nx=was_cfg["max_nx"], max_nt=nt_w,
                           dPML=40, fs=0, dmodel=dx, dt=dt_w, tile_x=16, snap_format="f16", store_snap=4,
                           kahan=True, hbm_budget=was_cfg["budget"])}
        return e

    rates = [float(r) for r in args.rates.split(",")]
    per_rate = []
    for ri, rate in enumerate(rates):
        arr = poisson_arrivals(rate, n, seed=10 + ri)
        reqs = [Request(q, float(arr[q]), pool_reqs[order[q]][0], pool_reqs[order[q]][1], kind=pool_reqs[order[q]][0])
                for q in range(n)]
        eng = engines()
        log(f"[bench] served: rate {rate}/s")
        res = Server(eng, fixed_bytes={k: e.memory()[0] for k, e in eng.items()}).run(reqs, log=log)
        for e in eng.values():
            e.close()
        res.pop("requests")
        res["rate_per_s"] = rate
        per_rate.append(res)
    best = max(per_rate, key=lambda r: r["shots_per_s"])
    out = {
        "metric": "migrated shots/s + Gpt-updates/s, batched vs plain, at rel-L2<=1e-4; p99 latency",
        "value": round(best["shots_per_s"], 4), "unit": "shots/s", "n_gpus": 1, "steps": n, "warmup": 0,
        "ms_per_step": round(1e3 / best["shots_per_s"], 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded models; seeded noise gathers; seeded Poisson arrivals)",
        "config": {"workload": "served stream: 70% MCS (480 ch @12.5 m, 8 km aperture, 601 deep @10 m, nt 6000, "
                               "dt 0.8 ms, FP32 snapshots) + 30% WAS (801x6001 @10 m, 50 km offset, nt 12000, "
                               "dt 1 ms, 8 Hz, FP16 snapshots every 4th step)", "form": "served",
                   "mcs_engine": {k: v for k, v in mcs_cfg.items()}, "was_engine": {k: v for k, v in was_cfg.items()}},
        "per_rate": per_rate,
        "p99_latency_s_at_best_rate": best["latency"]["p99_s"],
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def kv_cfg(text):
    """{"key": int} from "key=value,key=value" (None for None)."""
    if not text:
        return None
    return {k: int(v) for k, v in (kv.split("=") for kv in text.split(","))}


def bench_copy(args):
    """Calibration of the HBM traffic counters (rocprofv3 FETCH_SIZE / WRITE_SIZE): four
    device-to-device copies of 256 MiB, whose bytes read and written are known."""
    import torch

    dev = torch.device("cuda:0")
    a = torch.ones(64 * 1024 * 1024, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(4):
        b.copy_(a)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    gbs = 4 * 2 * a.numel() * 4 / el / 1e9
    print(json.dumps({"metric": "device copy GB/s (read + write)", "value": round(gbs, 1), "unit": "GB/s", "n_gpus": 1,
                      "steps": 4, "warmup": 0, "ms_per_step": round(1e3 * el / 4, 3), "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                      "config": {"workload": "copy: 4 x 256 MiB device to device (counter calibration)"}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64, help="shots timed (per form)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed shots before the plain timed region")
    ap.add_argument("--tile-x", type=int, default=8, help="plain form CPML tile width (8, 16 or 32)")
    ap.add_argument("--timing-every", type=int, default=16, help="sample every n-th step kernel with HIP events")
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="CPU baseline processes (0: every core this process may run on, at most 16, the host "
                         "cores one GPU of the box is allotted)")
    ap.add_argument("--cpu-steps", type=int, default=1200, help="CPU baseline: time steps of the sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-wide", action="store_true", help="skip the plain form on the 801 x 5102 window")
    ap.add_argument("--wide-steps", type=int, default=1500, help="plain form on the 801 x 5102 window: steps timed")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-zero-sum", action="store_true", help="skip the batched form with zero-sum weights")
    ap.add_argument("--no-tune", action="store_true", help="plain and batched forms with the default launch shapes")
    ap.add_argument("--plain-cfg", default=None,
                    help="plain form on the benchmark's window: launch shape tile_x=..,stream_w=.. instead of the "
                         "per-window tuning (to profile a tuned run's shape)")
    ap.add_argument("--wide-cfg", default=None,
                    help="plain form on the 801 x 5102 window: launch shape tile_x=..,stream_w=.. instead of tuning")
    ap.add_argument("--engine-cfg", default=None,
                    help="batched form: launch shape as key=value pairs, e.g. nslots=20,tile_x=16,stream_w=4,groups=2 "
                         "(keys nslots, tile_x, stream_w, grid, groups, steps_per_launch), used instead of the "
                         "autotuner's (to profile a tuned run's shape)")
    ap.add_argument("--synthetic-gathers", action="store_true",
                    help="mcs2001: seeded noise gathers instead of gathers modelled on the device (same cost per "
                         "migrated shot; keeps profiled runs short)")
    ap.add_argument("--forms", default="plain,batched")
    ap.add_argument("--workload", default="mcs2001",
                    choices=["mcs2001", "uneven", "mixed", "served", "condition", "poststack", "su", "kirchhoff",
                             "copy"],
                    help="mcs2001: the 64-shot line (default line); uneven: 256 uneven shots, batched vs plain; "
                         "mixed: uneven + FP16/BF16 snapshots every 4th step; served: Poisson-arrival stream; "
                         "condition / poststack / su: the steps before and after the migration; "
                         "kirchhoff: the Kirchhoff migration of the line's shots; copy: a 256 MiB device copy, "
                         "the calibration of the HBM traffic counters")
    ap.add_argument("--shots", type=int, default=256, help="uneven/mixed: shots of the 256-shot line used")
    ap.add_argument("--slots", type=int, default=16, help="uneven/mixed/served: engine slots")
    ap.add_argument("--hbm-gb", type=float, default=200.0, help="uneven/mixed/served: snapshot budget (GiB)")
    ap.add_argument("--no-plain", action="store_true", help="uneven/mixed: skip the plain form")
    ap.add_argument("--requests", type=int, default=40, help="served: requests per rate")
    ap.add_argument("--rates", default="1,4,8,16",
                    help="served: Poisson arrival rates (requests/s); the upper ones exceed the stream's capacity")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library runs the RTM path on a single MI355X "
              f"(--gpus {args.gpus} is not supported; shots shard across GPUs with no per-step "
              f"communication)", file=sys.stderr)
        sys.exit(2)
    if args.cpu_procs <= 0:  # every core this process may run on (at most 16), for every workload
        args.cpu_procs = host_cores()[0]
    if args.workload in ("uneven", "mixed"):
        return bench_uneven(args, mixed=args.workload == "mixed")
    if args.workload == "served":
        return bench_served(args)
    if args.workload == "condition":
        return bench_condition(args)
    if args.workload == "poststack":
        return bench_poststack(args)
    if args.workload == "su":
        return bench_su(args)
    if args.workload == "kirchhoff":
        return bench_kirchhoff(args)
    if args.workload == "copy":
        return bench_copy(args)
    forms = set(args.forms.split(","))

    from psdm_amd.workloads import line_workload, step_bytes, updates_per_shot

    W = line_workload("mcs2001")
    m, shots, nt, dt, dPML = W["model"], W["shots"], W["nt"], W["dt"], W["dPML"]
    ny = m.nz
    nxw = shots[0].nx_w

    cpu = None
    if not args.no_cpu:
        procs = args.cpu_procs
        log(f"[bench] cpu baseline: {procs} procs x {args.cpu_steps} steps")
        cpu = cpu_leg(ny, nxw, dPML, m.dx, dt, procs, args.cpu_steps, nt)
        log(f"[bench] cpu baseline {cpu['value']} Gupd/s ({cpu['kind']})")

    import numpy as np
    import torch

    from psdm_amd._native import Context
    from psdm_amd.solvers import DeviceShot, ricker_mcs

    dev = torch.device("cuda:0")
    ctx = Context(0)
    ctx.set_launch(tile_x=args.tile_x)
    d_true = torch.from_numpy(np.ascontiguousarray(m.true.T)).to(dev)  # (nx, ny) = Fortran (ny, nx)
    d_mig = torch.from_numpy(np.ascontiguousarray(m.migration.T)).to(dev)
    d_wat = torch.from_numpy(np.ascontiguousarray(m.water.T)).to(dev)
    src = ricker_mcs(nt, dt, W["f0"], 1.2 / W["f0"])
    d_src = torch.from_numpy(src.reshape(1, nt)).to(dev)
    torch.cuda.synchronize()

    nshots = min(max(args.steps + args.warmup, 1), len(shots))
    geos, geos_model = [], []
    for k in range(nshots):
        sh = shots[k]
        args_geo = (ny, sh.nx_w, nt, dPML, W["fs"], m.dx, dt, [sh.nxSou - sh.x_i + 1], [sh.nySou],
                    sh.nxRec - sh.x_i + 1, sh.nyRec)
        geos.append(DeviceShot(*args_geo))                   # migration: snapshots every step
        geos_model.append(DeviceShot(*args_geo, nt_snap=0))  # data modelling: no snapshots
    if args.synthetic_gathers:
        log(f"[bench] {nshots} seeded noise gathers")
        obs = random_gathers(torch, dev, shots[:nshots], [nt] * nshots, seed=0)
    else:
        log(f"[bench] modelling {nshots} gathers")
        obs = model_gathers(ctx, torch, dev, W, geos_model, d_true, d_wat, d_src, nshots)

    bfw = step_bytes(ny, nxw, dPML, 4.0, 1, backward=False)
    bbw = step_bytes(ny, nxw, dPML, 4.0, 1, backward=True)
    step_b = 0.5 * (bfw + bbw)  # algorithmic bytes of one shot-step (FP32 snapshot every step)
    upd_shot = updates_per_shot(ny, nxw, nt, dPML)

    plain = batched = None
    if "plain" in forms:
        P = run_plain(ctx, torch, dev, W, geos, d_mig, d_src, obs, args.steps, args.warmup, args.timing_every,
                      tune=not args.no_tune, shape=kv_cfg(args.plain_cfg))
        ach = step_b / (P["mean_ms"] * 1e-3) / 1e9 if P["mean_ms"] > 0 else None
        plain = {"shots_per_s": round(args.steps / P["elapsed"], 4),
                 "gpt_updates_per_s": round(args.steps * upd_shot / P["elapsed"] / 1e9, 3),
                 "ms_per_shot": round(1e3 * P["elapsed"] / args.steps, 3),
                 "roofline": {"bound": "hbm", "achieved": round(ach, 1) if ach else None, "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4) if ach else None,
                              "kernel": "rtm_step_kernel", "mean_kernel_ms": round(P["mean_ms"], 5),
                              "sampled_launches": P["n_ev"], "algorithmic_bytes_per_launch": round(step_b)},
                 "launch": P["launch"]}
        log(f"[bench] plain {plain['shots_per_s']} shots/s")
        if not args.no_wide:
            plain["wide_window"] = run_plain_wide(ctx, torch, dev, args.wide_steps, args.timing_every,
                                                  tune=not args.no_tune, shape=kv_cfg(args.wide_cfg))
            log(f"[bench] plain, wide window: roofline {plain['wide_window']['roofline']['frac']}")

    if "batched" in forms:
        cfg = dict(nslots=16, tile_x=16, stream_w=4, grid=0, groups=2, chunk_steps=64)
        tuned = None
        if args.engine_cfg:
            cfg.update({k: int(v) for k, v in (kv.split("=") for kv in args.engine_cfg.split(","))})
        elif not args.no_tune:
            from psdm_amd.autotune import tune

            jobs = [job_of(W, k, d_src, obs) for k in range(nshots)]
            best, trials = tune(ctx, d_mig, ny, m.nx, jobs, max(s.nx_w for s in shots), nt, dPML, W["fs"], m.dx, dt,
                                nshots=args.steps, steps_per_launch=(1, 2), log=log)
            cfg.update(nslots=best["nslots"], tile_x=best["tile_x"], stream_w=best["stream_w"], grid=best["grid"],
                       groups=best["groups"], steps_per_launch=best["steps_per_launch"])
            tuned = {"best": best, "trials": trials}
        B = run_batched(ctx, torch, dev, W, d_mig, d_src, obs, args.steps, args.warmup, cfg)
        st = B["stats"]
        spl = cfg.get("steps_per_launch", 1)
        # algorithmic bytes of one shot-step in the launch form that ran (two steps per launch
        # move the wavefield once per two steps)
        step_bb = 0.5 * (step_bytes(ny, nxw, dPML, 4.0, 1, backward=False, steps_per_launch=spl) +
                         step_bytes(ny, nxw, dPML, 4.0, 1, backward=True, steps_per_launch=spl))
        # device time per batched step (events around the captured chunks); one step is `groups`
        # concurrent dispatches of the batch kernel, one per slot group
        per_launch_ms = B["launch_ms"]
        shot_steps_launch = st["slot_steps"] / max(st["steps_issued"], 1)
        bytes_launch = step_bb * shot_steps_launch
        ach = bytes_launch / (per_launch_ms * 1e-3) / 1e9
        batched = {"shots_per_s": round(args.steps / B["elapsed"], 4),
                   "gpt_updates_per_s": round(args.steps * upd_shot / B["elapsed"] / 1e9, 3),
                   "config": cfg, "autotune": tuned,
                   "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "rtm_batch_kernel",
                                "mean_kernel_ms": round(per_launch_ms, 5), "timed_launches": B["n_launch"],
                                "concurrent_dispatches_per_launch": cfg["groups"],
                                "algorithmic_bytes_per_launch": round(bytes_launch),
                                "shot_steps_per_launch": round(shot_steps_launch, 4),
                                "wall_ms_per_launch": round(1e3 * B["elapsed"] / max(st["steps_issued"], 1), 5),
                                "steps_per_launch": spl,
                                "note": "one launch = steps_per_launch time steps of the resident shots, issued as "
                                        "concurrent_dispatches_per_launch dispatches (slot groups); mean_kernel_ms: HIP "
                                        "events around each captured chunk of launches / launches per chunk"}}
        if plain is not None:
            a = B["psdm_sum"].cpu().numpy()
            b = P["psdm_sum"].cpu().numpy()
            batched["stack_rel_diff_vs_plain"] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
            batched["fold_equal"] = bool(torch.equal(B["fold"], P["fold"]))
        log(f"[bench] batched {batched['shots_per_s']} shots/s")
        if not args.no_zero_sum:
            # the same shots and launch shape with the opt-in zero-sum weights (one step per launch)
            zcfg = dict(cfg, weights="zero_sum", steps_per_launch=1)
            BZ = run_batched(ctx, torch, dev, W, d_mig, d_src, obs, args.steps, args.warmup, zcfg)
            stz = BZ["stats"]
            zb = 0.5 * (step_bytes(ny, nxw, dPML, 4.0, 1, backward=False) +
                        step_bytes(ny, nxw, dPML, 4.0, 1, backward=True)) * stz["slot_steps"] / max(stz["steps_issued"], 1)
            achz = zb / (BZ["launch_ms"] * 1e-3) / 1e9
            batched["zero_sum"] = {
                "shots_per_s": round(args.steps / BZ["elapsed"], 4),
                "gpt_updates_per_s": round(args.steps * upd_shot / BZ["elapsed"] / 1e9, 3),
                "mean_kernel_ms": round(BZ["launch_ms"], 5), "roofline_frac": round(achz / HBM_PEAK_GBS, 4),
                "config": {k: v for k, v in zcfg.items()},
                "stack_rel_l2_vs_reference_weights": float(
                    np.linalg.norm((BZ["psdm_sum"] - B["psdm_sum"]).cpu().numpy()) /
                    max(np.linalg.norm(B["psdm_sum"].cpu().numpy()), 1e-300))}
            log(f"[bench] batched, zero-sum weights {batched['zero_sum']['shots_per_s']} shots/s")

    def traffic_of(name, key):
        """HBM bytes measured with rocprofv3 counters (profiles/<name>), per `key` unit."""
        import re

        tpath = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(tpath):
            return None
        mt = re.search(rf"^{key} = (\d+)", open(tpath).read(), re.M)
        return int(mt.group(1)) if mt else None

    acc = None
    if not args.no_accuracy and not args.synthetic_gathers:
        log("[bench] accuracy check vs the FP64 oracle fixtures")
        acc = {"short": accuracy_check(ctx)}
        if 0 in obs:
            acc["full"] = accuracy_line(ctx, torch, dev, W, geos, d_mig, d_src, obs,
                                        batched["config"] if batched else dict(tile_x=16, stream_w=4))
            log(f"[bench] full-record accuracy: plain {acc['full']['plain']['image_rel_l2_vs_fp64']:.3g}, "
                f"batched {acc['full']['batched']['image_rel_l2_vs_fp64']:.3g}, zero-sum weights "
                f"{acc['full']['zero_sum']['batched']['image_rel_l2_vs_fp64']:.3g} (bar {ACC_BAR})")

    head = batched or plain
    form = "batched" if batched else "plain"
    if plain is not None:
        plain["roofline"]["traffic"] = traffic_of("traffic_rtm_step_kernel_mcs2001.md", "traffic_bytes_per_launch")
    if batched is not None:
        # counters give bytes per shot-step (one launch's bytes / its busy slots): scaled to the
        # shots of the launches timed here, like algorithmic_bytes_per_launch
        per_shot_step = traffic_of("traffic_rtm_batch_kernel_mcs2001.md", "traffic_bytes_per_shot_step")
        spl = batched["roofline"]["shot_steps_per_launch"]
        batched["roofline"]["traffic"] = round(per_shot_step * spl) if per_shot_step else None
        batched["roofline"]["traffic_per_shot_step"] = per_shot_step
    roof = dict(head["roofline"])
    out = {
        "metric": "migrated shots/s + Gpt-updates/s, batched vs plain, at rel-L2<=1e-4; p99 latency",
        "value": head["shots_per_s"],
        "unit": "shots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / head["shots_per_s"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": ("synthetic (seeded 2001x601 model; seeded noise gathers)" if args.synthetic_gathers else
                 "synthetic (seeded 2001x601 model; gathers modelled on device, direct wave removed)"),
        "config": {"workload": (f"mcs2001 line: 2001x601 @10 m, 64 shots @25 m, 480 ch @12.5 m, near offset 150 m, "
                                f"dt 0.8 ms, nt 6000, dPML 40, FP32 snapshots every step; timed: {args.steps} shots "
                                f"(shots {args.warmup % len(shots)}..{(args.warmup + args.steps - 1) % len(shots)} of "
                                f"the line, after {args.warmup} untimed warmup shot(s) in the plain form)"),
                   "form": form, "window": f"{ny}x{nxw}"},
        "gpt_updates_per_s": head["gpt_updates_per_s"],
        "roofline": roof,
        "plain": plain,
        "batched": batched,
        "speedup_batched_vs_plain": (round(batched["shots_per_s"] / plain["shots_per_s"], 3)
                                     if batched and plain else None),
        "cpu_baseline": cpu,
        "accuracy": acc,
    }
    if batched is not None and "zero_sum" in batched and acc is not None and "full" in acc:
        z = acc["full"]["zero_sum"]
        out["within_accuracy_bar"] = {
            "bar": ACC_BAR, "reference_weights": acc["full"]["within_bar"], "zero_sum_weights": z["within_bar"],
            "shots_per_s_zero_sum_batched": batched["zero_sum"]["shots_per_s"],
            "image_rel_l2_vs_fp64": {"reference_weights": acc["full"]["batched"]["image_rel_l2_vs_fp64"],
                                     "zero_sum_weights": z["batched"]["image_rel_l2_vs_fp64"]},
            "note": "value is the reference-weight form (bit-exact with the reference's FP32 solver); the opt-in "
                    "zero-sum form is the one within the FP64 bar over the full record"}
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
