"""Seeded synthetic marine workloads for the RTM path (host side, numpy).

Velocity model: a 1500 m/s water layer over a sediment gradient (1800 -> 3500 m/s),
three dipping reflectors (+300 / -300 / +300 m/s steps, dips 0 / 5 / 15 degrees) and one
Gaussian high-velocity anomaly (sigma 300 m).  The migration model is the true model
smoothed with a 250 m Gaussian.  Acquisition follows the reference's MCS geometry rules
(get_mcs_data.f90:137-144, :500-501): nxSou = 1 + floor(x_shot/dx),
nySou = 1 + floor(z_shot/dx), pos_rec = x_shot - dir_boat*(near_offset + (irec-1)*drec).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Model:
    nz: int
    nx: int
    dx: float
    true: np.ndarray       # (nz, nx) float32, Fortran order (Vpmodel)
    migration: np.ndarray  # (nz, nx) float32, Fortran order
    water: np.ndarray      # (nz, nx) float32, water velocity everywhere (direct-wave modelling)
    water_depth: float
    reflectors: list = field(default_factory=list)  # (z at x=0, dip deg, dv)


def _gauss_smooth(a: np.ndarray, sigma_cells: float) -> np.ndarray:
    from scipy.ndimage import gaussian_filter

    return gaussian_filter(a.astype(np.float64), sigma_cells, mode="nearest")


def velocity_model(nz: int, nx: int, dx: float, water_depth: float, seed: int = 0) -> Model:
    rng = np.random.default_rng(seed)
    z = (np.arange(nz) * dx)[:, None]
    x = (np.arange(nx) * dx)[None, :]
    zmax = (nz - 1) * dx
    xmax = (nx - 1) * dx
    v = np.where(z < water_depth, 1500.0, 1800.0 + 1700.0 * (z - water_depth) / max(zmax - water_depth, dx))
    v = np.broadcast_to(v, (nz, nx)).copy()
    sed = zmax - water_depth
    reflectors = []
    for frac, dip, dv in ((0.25, 0.0, 300.0), (0.5, 5.0, -300.0), (0.75, 15.0, 300.0)):
        z0 = water_depth + sed * (frac + 0.04 * (rng.random() - 0.5))
        zr = z0 + np.tan(np.radians(dip)) * (x - 0.5 * xmax)
        v = v + np.where((z >= zr) & (z >= water_depth), dv, 0.0)
        reflectors.append((float(z0 - np.tan(np.radians(dip)) * 0.5 * xmax), dip, dv))
    xa = xmax * (0.35 + 0.3 * rng.random())
    za = water_depth + sed * (0.45 + 0.1 * rng.random())
    anomaly = 500.0 * np.exp(-((x - xa) ** 2 + (z - za) ** 2) / (2 * 300.0 ** 2))
    v = v + np.where(z >= water_depth, anomaly, 0.0)
    true = np.asfortranarray(v.astype(np.float32))
    mig = np.asfortranarray(_gauss_smooth(v, 250.0 / dx).astype(np.float32))
    water = np.asfortranarray(np.full((nz, nx), 1500.0, np.float32))
    return Model(nz, nx, dx, true, mig, water, water_depth, reflectors)


@dataclass
class Shot:
    index: int
    x_shot: float
    nxSou: int
    nySou: int
    nxRec: np.ndarray  # 1-based model columns
    nyRec: np.ndarray
    # aperture window (dynamic_window.f90:1-52)
    nx_w: int = 0
    x_i: int = 0
    x_f: int = 0


def mcs_line(model_nx: int, dx: float, nshots: int, first_x: float, shot_spacing: float, nrec: int, drec: float,
             near_offset: float, shot_depth: float, streamer_depth: float, dir_boat: int = -1,
             added_space: float = 500.0):
    from .solvers import dynamic_window

    shots = []
    for s in range(nshots):
        xs = first_x + s * shot_spacing
        pos_rec = xs - dir_boat * (near_offset + np.arange(nrec) * drec)
        nxR = (1 + np.floor(pos_rec / dx)).astype(np.int32)
        nyR = np.full(nrec, 1 + int(np.floor(streamer_depth / dx)), np.int32)
        sh = Shot(s, xs, 1 + int(np.floor(xs / dx)), 1 + int(np.floor(shot_depth / dx)), nxR, nyR)
        sh.nx_w, sh.x_i, sh.x_f = dynamic_window([sh.nxSou], nxR, dx, model_nx, added_space)
        shots.append(sh)
    return shots


def line_workload(name: str):
    """The two named line workloads: 'mcs2001' (2001 x 601 @ 10 m, dt 0.8 ms, nt 6000,
    64 shots at 25 m, 480 channels at 12.5 m from 150 m offset) and 'single401'
    (401 x 201 @ 12.5 m, dt 1 ms, nt 3000, one shot at 500 m, 240 channels at 12.5 m
    from 100 m offset).  Source / streamer depths 6 m / 8 m."""
    if name == "mcs2001":
        m = velocity_model(601, 2001, 10.0, 1500.0, seed=0)
        shots = mcs_line(2001, 10.0, 64, 1000.0, 25.0, 480, 12.5, 150.0, 6.0, 8.0)
        return dict(model=m, shots=shots, dt=0.0008, nt=6000, f0=15.0, dPML=40, fs=0)
    if name == "single401":
        m = velocity_model(201, 401, 12.5, 1000.0, seed=0)
        shots = mcs_line(401, 12.5, 1, 500.0, 0.0, 240, 12.5, 100.0, 6.0, 8.0)
        return dict(model=m, shots=shots, dt=0.001, nt=3000, f0=15.0, dPML=40, fs=0)
    raise ValueError(name)


def updates_per_shot(ny: int, nx_w: int, nt: int, dPML: int) -> int:
    """Padded grid-point updates of one migrated shot (forward + backward)."""
    dn = dPML + 6
    return 2 * (ny + 2 * dn) * (nx_w + 2 * dn) * (nt - 1)


def step_bytes(ny: int, nx_w: int, dPML: int, snap_bytes: float = 4.0, snap_every: int = 1, backward=False,
               steps_per_launch: int = 1):
    """Algorithmic HBM bytes of one time step: 16 B per padded point (read p, p_prev,
    c2, write p), 16 B per strip point and 32 B per corner point (CPML phi/chi read +
    write per direction), plus the snapshot write (forward) or snapshot read + image
    read-modify-write (backward) over the interior every `snap_every` steps.

    With two steps per launch the wavefield and the memories cross HBM once per launch:
    per two steps 20 B per padded point (read p(k-1), p(k-2), c2; write p(k+1), p(k)),
    16 B per strip point and 32 B per corner point, the image read-modify-write once per
    launch that images; the snapshots stay per step."""
    dn = dPML + 6
    n_pad = (ny + 2 * dn) * (nx_w + 2 * dn)
    n_strip = 2 * dn * ny + 2 * dn * nx_w
    n_corner = 4 * dn * dn
    n_int = ny * nx_w
    if steps_per_launch == 2:
        img = 8.0 / max(2, snap_every) if backward else 0.0  # once per launch that images
        return (20.0 * n_pad + 16.0 * n_strip + 32.0 * n_corner) / 2 + (snap_bytes / snap_every + img) * n_int
    per = (snap_bytes + (8.0 if backward else 0.0)) / snap_every
    return 16.0 * n_pad + 16.0 * n_strip + 32.0 * n_corner + per * n_int


def _mcs_shot(index, xs, dx, model_nx, nrec, drec, near_offset, shot_depth, streamer_depth, added_space,
              dir_boat=-1):
    from .solvers import dynamic_window

    pos_rec = xs - dir_boat * (near_offset + np.arange(nrec) * drec)
    nxR = (1 + np.floor(pos_rec / dx)).astype(np.int32)
    nyR = np.full(nrec, 1 + int(np.floor(streamer_depth / dx)), np.int32)
    sh = Shot(index, xs, 1 + int(np.floor(xs / dx)), 1 + int(np.floor(shot_depth / dx)), nxR, nyR)
    sh.nx_w, sh.x_i, sh.x_f = dynamic_window([sh.nxSou], nxR, dx, model_nx, added_space)
    return sh


def uneven_line(nshots: int = 256, seed: int = 0, line_km: float = 40.0, nz: int = 601, dx: float = 10.0,
                nx_range=(600, 2000), nts=(4000, 6000, 8000)):
    """Shots along a `line_km` line whose aperture windows are uneven: each shot's window
    width is drawn uniformly from `nx_range` columns (the streamer is sized to it: 12.5 m
    channels from 150 m near offset, 500 m added space each side) and its record length
    from `nts` (dt 0.8 ms).  Returns dict(model, shots, nt per shot, dt, f0, dPML, fs)."""
    rng = np.random.default_rng(seed)
    model_nx = int(round(line_km * 1000.0 / dx)) + 1
    m = velocity_model(nz, model_nx, dx, 1500.0, seed=seed)
    drec, near, added = 12.5, 150.0, 500.0
    ddx = 1 + int(np.floor(added / dx))
    shots, nt = [], []
    for s in range(nshots):
        w = int(rng.integers(nx_range[0], nx_range[1] + 1))
        span_cols = max(w - 2 * ddx - int(near / dx), 2)
        nrec = max(2, int(np.floor(span_cols * dx / drec)) + 1)
        reach = near + (nrec - 1) * drec + added
        xs = float(rng.uniform(added + dx, (model_nx - 1) * dx - reach - dx))
        shots.append(_mcs_shot(s, xs, dx, model_nx, nrec, drec, near, 6.0, 8.0, added))
        nt.append(int(rng.choice(nts)))
    return dict(model=m, shots=shots, nt=nt, dt=0.0008, f0=15.0, dPML=40, fs=0)


def was_gathers(n: int, seed: int = 0, nz: int = 801, nx: int = 6001, dx: float = 10.0, offset_km: float = 50.0,
                shot_spacing: float = 50.0):
    """Long-offset wide-angle gathers treated by reciprocity: a seafloor receiver (OBS)
    acts as the source and the shots of one source line (6 m depth, every
    `shot_spacing` m out to `offset_km`) act as receivers; grid nz x nx at dx,
    dt 1 ms, nt 12000, 8 Hz Ricker.  OBS positions are drawn so the line fits the model."""
    rng = np.random.default_rng(seed)
    m = velocity_model(nz, nx, dx, 1500.0, seed=seed)
    nrec = int(offset_km * 1000.0 / shot_spacing)
    out = []
    for s in range(n):
        x_obs = float(rng.uniform(600.0, (nx - 1) * dx - offset_km * 1000.0 - 600.0))
        sh = _mcs_shot(s, x_obs, dx, nx, nrec, shot_spacing, shot_spacing, 1500.0, 6.0, 500.0)
        out.append(sh)
    return dict(model=m, shots=out, nt=12000, dt=0.001, f0=8.0, dPML=40, fs=0)
