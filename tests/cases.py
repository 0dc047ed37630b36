"""Seeded synthetic shot configurations shared by the parity tests."""
from __future__ import annotations

import numpy as np


def ricker(nt, dt, f0, t0):
    t = np.arange(nt) * dt
    a = (np.pi * f0 * (t - t0)) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def make_case(ny=40, nx=50, nt=300, dPML=10, dx=10.0, dt=0.001, fs=0, variant=0, store_snap=1, nt_snap=None,
              seed=0, src=(20, 5), rec_x=None, rec_y=3, f0=25.0, vmin=1500.0, vmax=2500.0, extra_src=()):
    """Random smooth-ish velocity, a Ricker at `src` (nxSou, nySou) 1-based, receivers on a row."""
    rng = np.random.default_rng(seed)
    base = vmin + (vmax - vmin) * rng.random((ny, nx))
    # light smoothing keeps the model realistic while staying random
    k = np.ones(3) / 3
    base = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), 0, base)
    base = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), 1, base)
    ctmp = np.asfortranarray(base.astype(np.float32))
    S = ricker(nt, dt, f0, 1.2 / f0).reshape(nt, 1)
    nxSou = [src[0]] + [e[0] for e in extra_src]
    nySou = [src[1]] + [e[1] for e in extra_src]
    if extra_src:
        S = np.asfortranarray(np.concatenate([S] + [0.5 * S] * len(extra_src), axis=1))
    if rec_x is None:
        rec_x = np.arange(3, nx - 2, 3)
    rec_x = np.asarray(rec_x, dtype=np.int32)
    rec_y = np.broadcast_to(np.asarray(rec_y, dtype=np.int32), rec_x.shape).copy()
    return dict(ny=ny, nx=nx, nt=nt, dPML=dPML, dx=dx, dt=dt, fs=fs, variant=variant, store_snap=store_snap,
                nt_snap=nt if nt_snap is None else nt_snap, ctmp=ctmp, S=np.asfortranarray(S),
                nxSou=np.asarray(nxSou, np.int32), nySou=np.asarray(nySou, np.int32), nxRec=rec_x, nyRec=rec_y)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


# Cases exercised by the golden vectors and the GPU parity tests.  Together they cover
# both stepping variants, free surface on/off, decimated snapshots (with the
# reference's own nt_snap bookkeeping), duplicate receivers (last writer wins), points
# outside the interior, several tiles per axis and interior-only tiles.
CASES = {
    "ram_fs0": dict(),
    "ram_fs1": dict(fs=1, seed=1),
    "disk_fs1": dict(fs=1, variant=1, seed=2),
    "disk_fs0": dict(fs=0, variant=1, seed=3, src=(10, 20)),
    "ram_ss2": dict(store_snap=2, seed=4, nt=301, nt_snap=151),
    # the reference drivers pass nt_snap = nt: with store_snap > 1 the backward then
    # reads only snapshot slots the forward never filled, and the image is zero
    "ram_ss3_driver_quirk": dict(store_snap=3, seed=8, nt=301),
    "ram_dup_pml_rec": dict(seed=5, rec_x=[5, 5, 9, 9, 9, 14, -3, 60, 25], rec_y=[3, 3, 1, 1, 1, 40, 2, 5, -4],
                            extra_src=((20, 5), (3, 0))),
    "ram_big": dict(ny=140, nx=96, nt=220, seed=6, src=(48, 4), rec_y=2),
}


def case(name, **over):
    kw = dict(CASES[name])
    kw.update(over)
    return make_case(**kw)
