"""CPU baseline timing for bench.py -- TEST INFRASTRUCTURE ONLY (the checker's side).

Times the reference Fortran solvers (oracle/_ref/libpsdm_ref.so, kind "reference") or,
where that library was not built, the C restatement (oracle/liboracle.so, kind "port")
on a bounded sample of a shot: forward + backward over `nt_sample` time steps of the
shot's full padded window.  N independent processes each migrate their own copy, the
way `mpirun -np N psdm_run` gives each rank its own shot (calculate_PSDM_RTM.f90:151-168).
"""
from __future__ import annotations

import multiprocessing as mp
import os
import time

import numpy as np


def _one_rank(args):
    ny, nx, nt, dPML, dx, dt, kind, seed = args
    import sys

    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.dirname(here))
    from oracle import pyoracle as O

    rng = np.random.default_rng(seed)
    ctmp = np.asfortranarray((1500 + 2000 * rng.random((ny, nx))).astype(np.float32))
    t = np.arange(nt) * dt
    a = (np.pi * 15.0 * (t - 0.08)) ** 2
    S = ((1 - 2 * a) * np.exp(-a)).astype(np.float32).reshape(nt, 1)
    nxR = np.arange(10, nx - 10, 2, dtype=np.int32)
    nyR = np.ones_like(nxR)
    impl = "ref" if kind == "reference" else "oracle"
    t0 = time.perf_counter()
    Rec, Fields = O.forward([nx // 2], [1], nxR, nyR, S, dx, dt, ctmp, 1, nt, nt, dPML, 0, impl=impl)
    Sadj = np.asfortranarray(Rec[::-1, :])
    O.backward([nx // 2], [1], nxR, nyR, Sadj, dx, dt, ctmp, 1, nt, nt, dPML, 0, Fields=Fields, impl=impl)
    return time.perf_counter() - t0


def run(ny: int, nx: int, nt_sample: int, dPML: int, dx: float, dt: float, nprocs: int):
    from oracle import pyoracle as O

    kind = "reference" if O.ref_available() else "port"
    dn = dPML + 6
    updates = 2 * (ny + 2 * dn) * (nx + 2 * dn) * (nt_sample - 1)
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(nprocs) as pool:
        per = pool.map(_one_rank, [(ny, nx, nt_sample, dPML, dx, dt, kind, s) for s in range(nprocs)])
    wall = time.perf_counter() - t0
    busy = max(per)
    return dict(kind=kind, cores=nprocs, updates_per_rank=updates, wall_s=wall, rank_s=busy,
                gupd_per_s=nprocs * updates / busy / 1e9)


def _one_rank_conditioning(args):
    nt_data, dt_data, nrec, nt_prop2, dt_prop, typef, f1, f2, N_taper, seed = args
    import sys

    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.dirname(here))
    from oracle import conditioning as C

    rng = np.random.default_rng(seed)
    raw = np.asfortranarray(rng.standard_normal((nt_data, nrec)).astype(np.float32))
    t0 = time.perf_counter()
    f = C.ref_freq_filter(raw, dt_data, typef, f1, f2)
    out = C.ref_akima(f, dt_data, nt_prop2, dt_prop)
    for i in range(nrec):
        out[:, i] = C.ref_taper_pad_trace(out[:, i], nt_prop2, N_taper, 0)
    return time.perf_counter() - t0


def run_conditioning(nt_data: int, dt_data: float, nrec: int, nt_prop2: int, dt_prop: float, typef: int, f1: float,
                     f2: float, N_taper: int, nprocs: int):
    """Gather conditioning by the reference's own routines (oracle/_ref: freq_filter with
    FFTPACK, interpolation_akima, taper_pad_trace), one gather of `nrec` traces per
    process; returns traces per second over the slowest process."""
    from oracle import pyoracle as O

    if not O.ref_available():
        raise RuntimeError("oracle/_ref/libpsdm_ref.so is not built")
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(nprocs) as pool:
        per = pool.map(_one_rank_conditioning, [(nt_data, dt_data, nrec, nt_prop2, dt_prop, typef, f1, f2, N_taper, s)
                                                for s in range(nprocs)])
    wall = time.perf_counter() - t0
    busy = max(per)
    return dict(kind="reference", cores=nprocs, traces_per_rank=nrec, wall_s=wall, rank_s=busy,
                traces_per_s=nprocs * nrec / busy)


def _one_rank_twt(args):
    nz, ncols, dd, dt, seed = args
    import sys

    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.dirname(here))
    from oracle import poststack as P

    rng = np.random.default_rng(seed)
    z = np.arange(nz)[:, None] * dd
    vel = np.where(z < 1500.0, 1500.0, 1800.0 + 0.3 * (z - 1500.0)).astype(np.float32) * np.ones((1, ncols), np.float32)
    img = rng.standard_normal((nz, ncols)).astype(np.float32)
    t0 = time.perf_counter()
    t = P.twt_times(vel, dd)
    nt = P.twt_length(t, dt)
    time_axis = np.arange(nt).astype(np.float32) * np.float32(dt)
    for i in range(ncols):
        P.ref_linear_interpolation_1d(t[:, i], vel[:, i], time_axis)
        P.ref_linear_interpolation_1d(t[:, i], img[:, i], time_axis)
    return time.perf_counter() - t0


def run_twt(nz: int, ncols_per_rank: int, dd: float, dt: float, nprocs: int):
    """Depth-to-time conversion with the reference's linear_interpolation_1D (oracle/_ref),
    `ncols_per_rank` image columns per process; returns columns per second."""
    from oracle import pyoracle as O

    if not O.ref_available():
        raise RuntimeError("oracle/_ref/libpsdm_ref.so is not built")
    ctx = mp.get_context("spawn")
    with ctx.Pool(nprocs) as pool:
        per = pool.map(_one_rank_twt, [(nz, ncols_per_rank, dd, dt, s) for s in range(nprocs)])
    busy = max(per)
    return dict(kind="reference", cores=nprocs, cols_per_rank=ncols_per_rank, rank_s=busy,
                cols_per_s=nprocs * ncols_per_rank / busy)


def _one_rank_su(args):
    nrec, ns, reps, seed = args
    import tempfile

    from oracle import su_codec as S

    rng = np.random.default_rng(seed)
    buf = b"".join(rng.integers(0, 1000, 60).astype(">i4").tobytes() + rng.standard_normal(ns).astype(">f4").tobytes()
                   for _ in range(nrec))
    fd, path = tempfile.mkstemp(suffix=".su")
    try:
        with os.fdopen(fd, "wb") as f:
            f.write(buf)
        S.ref_read_shot(path, True, nrec, ns, ns)  # page cache warm
        t0 = time.perf_counter()
        for _ in range(reps):
            S.ref_read_shot(path, True, nrec, ns, ns)
        return time.perf_counter() - t0
    finally:
        os.unlink(path)


def run_su(nrec: int, ns: int, reps: int, nprocs: int):
    """SU gather reads with the reference's open_shot_mcs + get_shot_data (oracle/_ref,
    big-endian file in the page cache), `reps` reads per process; returns gathers per second."""
    from oracle import pyoracle as O

    if not O.ref_available():
        raise RuntimeError("oracle/_ref/libpsdm_ref.so is not built")
    ctx = mp.get_context("spawn")
    with ctx.Pool(nprocs) as pool:
        per = pool.map(_one_rank_su, [(nrec, ns, reps, s) for s in range(nprocs)])
    busy = max(per)
    return dict(kind="reference", cores=nprocs, reps=reps, rank_s=busy, gathers_per_s=nprocs * reps / busy)


def _one_rank_kirchhoff(args):
    ny, nx, dd, n_eik, n_img, nt, dt, seed = args
    from oracle import kirchhoff as K

    rng = np.random.default_rng(seed)
    z = np.arange(ny)[None, :]
    v = (np.where(z < 100, 1500.0, 1700.0 + 2.0 * z) + 100 * rng.random((nx, ny))).astype(np.float32)
    t0 = time.perf_counter()
    tts = K.eikonal(v, dd, dd, nx // 2, 1)
    ttr = np.stack([K.eikonal(v, dd, dd, 1 + (nx - 1) * q // max(1, n_eik - 1), 2) for q in range(n_eik - 1)])
    t_eik = (time.perf_counter() - t0) / n_eik
    data = rng.standard_normal((n_img, nt)).astype(np.float32)
    t1 = time.perf_counter()
    K.image(tts, ttr[np.arange(n_img) % ttr.shape[0]], data, nt, dd, dt)
    t_img = (time.perf_counter() - t1) / n_img
    return t_eik, t_img


def run_kirchhoff(ny: int, nx: int, dd: float, nrec: int, nt: int, dt: float, n_eik: int, n_img: int, nprocs: int):
    """Kirchhoff migration of one shot with the oracle's C restatement (the eikonal is the
    reference's solve_eikonal_FSM bit for bit): each process times n_eik travel-time solves
    and the imaging of n_img receivers; a shot is 1 + nrec solves and nrec receivers."""
    ctx = mp.get_context("spawn")
    with ctx.Pool(nprocs) as pool:
        per = pool.map(_one_rank_kirchhoff, [(ny, nx, dd, n_eik, n_img, nt, dt, s) for s in range(nprocs)])
    shot_s = max((1 + nrec) * te + nrec * ti for te, ti in per)
    return dict(kind="port", cores=nprocs, eikonal_s=max(te for te, _ in per), image_s_per_rec=max(ti for _, ti in per),
                shot_s=shot_s, shots_per_s=nprocs / shot_s)
