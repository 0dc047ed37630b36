"""Host-side mirror of PSDM's RTM solver seam, backed by libpsdm_rtm.so.

The four solver entry points keep the reference's names, argument order and meaning
(solver_ac_forward_ram.f90:4-6, solver_ac_backward_ram.f90:4-6,
solver_ac_forward_disk.f90:4-7, solver_ac_backward_disk.f90:4-7): Fortran-ordered
arrays, 1-based geometry, outputs filled in place (``Rec_point`` overwritten,
``kernel`` accumulated).  The module switches the solvers read from ``mod_parfile``
(dPML, fs and, for the snapshot hand-over, taper_option) come from a :class:`ParFile`.

``dynamic_window`` and the ``dynamic_window_solver_*`` wrappers mirror
dynamic_window.f90:1-228, and ``ricker_mcs`` mirrors get_mcs_data.f90:318-351.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from ._native import Context, Geom

_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


@dataclass
class ParFile:
    """The mod_parfile values on the solver path, with the reference defaults
    (modules.f90:195 dPML=20, :209 free_surface=1; taper_option 0)."""

    dPML: int = 20
    fs: int = 1
    taper_option: int = 0


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).ravel())


def _f32F(a, shape=None) -> np.ndarray:
    arr = np.asarray(a, dtype=np.float32)
    if shape is not None:
        arr = arr.reshape(shape, order="F")
    return np.asfortranarray(arr)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _check_out(arr, shape, name):
    if not (isinstance(arr, np.ndarray) and arr.dtype == np.float32 and arr.flags.f_contiguous):
        raise TypeError(f"{name} must be a Fortran-ordered float32 array")
    if arr.shape != tuple(shape):
        raise ValueError(f"{name} has shape {arr.shape}, expected {tuple(shape)}")


def solver_ac_forward_ram(rank, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, S, dmodel, dt, ny, nx, ctmp,
                          store_snap, nt_snap, nt, Fields, Rec_point, *, par: ParFile = ParFile(),
                          host_fields: Optional[bool] = None, ctx: Optional[Context] = None):
    """solver_ac_forward_ram (solver_ac_forward_ram.f90:4-6).  Rec_point(nt,NumRec) is
    overwritten.  Snapshots go to Fields(ny,nx,nt_snap) when ``host_fields`` (default:
    par.taper_option != 0, as the driver then tapers them, calculate_PSDM_RTM.f90:421-427);
    otherwise they stay in HBM keyed by Fields for solver_ac_backward_ram."""
    ctx = ctx or default_context()
    if host_fields is None:
        host_fields = par.taper_option != 0
    _check_out(Rec_point, (nt, NumRec), "Rec_point")
    _check_out(Fields, (ny, nx, nt_snap), "Fields")
    nxS, nyS, nxR, nyR = _i32(nxSou), _i32(nySou), _i32(nxRec), _i32(nyRec)
    S_ = _f32F(S, (nt, NumSou))
    c_ = _f32F(ctmp, (ny, nx))
    rc = ctx._L.psdm_rtm_forward_ram(ctx.h, par.dPML, par.fs, NumSou, NumRec, _ptr(nxS), _ptr(nyS), _ptr(nxR),
                                     _ptr(nyR), _ptr(S_), dmodel, dt, ny, nx, _ptr(c_), store_snap, nt_snap, nt,
                                     _ptr(Fields), int(host_fields), _ptr(Rec_point))
    ctx.check(rc, "solver_ac_forward_ram")


def solver_ac_backward_ram(rank, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, Sadj, dmodel, dt, ny, nx, ctmp,
                           store_snap, nt_snap, nt, Fields, kernel, *, par: ParFile = ParFile(),
                           host_fields: Optional[bool] = None, ctx: Optional[Context] = None):
    """solver_ac_backward_ram (solver_ac_backward_ram.f90:4-6).  kernel(ny,nx) += image."""
    ctx = ctx or default_context()
    if host_fields is None:
        host_fields = par.taper_option != 0
    _check_out(kernel, (ny, nx), "kernel")
    if host_fields:
        Fields = _f32F(Fields, (ny, nx, nt_snap))
    nxS, nyS, nxR, nyR = _i32(nxSou), _i32(nySou), _i32(nxRec), _i32(nyRec)
    S_ = _f32F(Sadj, (nt, NumRec))
    c_ = _f32F(ctmp, (ny, nx))
    rc = ctx._L.psdm_rtm_backward_ram(ctx.h, par.dPML, par.fs, NumSou, NumRec, _ptr(nxS), _ptr(nyS), _ptr(nxR),
                                      _ptr(nyR), _ptr(S_), dmodel, dt, ny, nx, _ptr(c_), store_snap, nt_snap, nt,
                                      _ptr(Fields), int(host_fields), _ptr(kernel))
    ctx.check(rc, "solver_ac_backward_ram")


def solver_ac_forward_disk(rank, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, S, dmodel, dt, ny, nx, ctmp, shotid,
                           unit_fields, folder_fields, store_snap, nt_snap, nt, Rec_point, *,
                           par: ParFile = ParFile(), ctx: Optional[Context] = None):
    """solver_ac_forward_disk (solver_ac_forward_disk.f90:4-7).  The per-step snapshot
    files become device-resident snapshots keyed by shotid; unit_fields and
    folder_fields are accepted for signature compatibility and unused."""
    ctx = ctx or default_context()
    _check_out(Rec_point, (nt, NumRec), "Rec_point")
    nxS, nyS, nxR, nyR = _i32(nxSou), _i32(nySou), _i32(nxRec), _i32(nyRec)
    S_ = _f32F(S, (nt, NumSou))
    c_ = _f32F(ctmp, (ny, nx))
    rc = ctx._L.psdm_rtm_forward_disk(ctx.h, par.dPML, par.fs, NumSou, NumRec, _ptr(nxS), _ptr(nyS), _ptr(nxR),
                                      _ptr(nyR), _ptr(S_), dmodel, dt, ny, nx, _ptr(c_), shotid, store_snap,
                                      nt_snap, nt, _ptr(Rec_point))
    ctx.check(rc, "solver_ac_forward_disk")


def solver_ac_backward_disk(rank, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, Sadj, dmodel, dt, ny, nx, ctmp, shotid,
                            unit_fields, folder_fields, store_snap, nt_snap, nt, kernel, *,
                            par: ParFile = ParFile(), ctx: Optional[Context] = None):
    """solver_ac_backward_disk (solver_ac_backward_disk.f90:4-7); consumes the snapshots
    solver_ac_forward_disk left under shotid."""
    ctx = ctx or default_context()
    _check_out(kernel, (ny, nx), "kernel")
    nxS, nyS, nxR, nyR = _i32(nxSou), _i32(nySou), _i32(nxRec), _i32(nyRec)
    S_ = _f32F(Sadj, (nt, NumRec))
    c_ = _f32F(ctmp, (ny, nx))
    rc = ctx._L.psdm_rtm_backward_disk(ctx.h, par.dPML, par.fs, NumSou, NumRec, _ptr(nxS), _ptr(nyS), _ptr(nxR),
                                       _ptr(nyR), _ptr(S_), dmodel, dt, ny, nx, _ptr(c_), shotid, store_snap,
                                       nt_snap, nt, _ptr(kernel))
    ctx.check(rc, "solver_ac_backward_disk")


# ----------------------------------------------------------------------------------
# Aperture window (dynamic_window.f90:1-52) and the solver wrappers (:54-228)
# ----------------------------------------------------------------------------------

def dynamic_window(nxSou, nxRec, dd: float, nx: int, dd_space: float):
    """Column window [x_i, x_f] (1-based, inclusive) of one shot: the source and
    receiver columns widened by ddx = 1 + floor(dd_space/dd) cells on each side,
    clipped to the model (dynamic_window.f90:15-50).  Returns (nx_w, x_i, x_f)."""
    ddx = 1 + int(np.floor(np.float32(dd_space) / np.float32(dd)))
    nxs = _i32(nxSou)
    nxr = _i32(nxRec)
    dmin = int(min(nxs.min(), nxr.min()))
    dmax = int(max(nxs.max(), nxr.max()))
    x_i = max(dmin - ddx + 1, 1)
    x_f = min(dmax + ddx, nx)
    return x_f - x_i + 1, x_i, x_f


def ricker_mcs(nt_prop: int, dt_prop: float, freq_ricker: float, delay_ricker: float = 0.0) -> np.ndarray:
    """Ricker_mcs (get_mcs_data.f90:318-351) in FP32: S0 = (1-2a)exp(-a),
    a = (pi f0 (t-t0))^2, t = (i-1) dt, zero for t >= 10/f0.  exp is evaluated in FP64
    and rounded once to FP32 (the correctly rounded FP32 exponential, as the Fortran
    runtime's expf returns it)."""
    f32 = np.float32
    pi, f0, t0 = f32(3.14159265), f32(freq_ricker), f32(delay_ricker)
    tmp = f32(10.0) / f0
    t = np.arange(nt_prop, dtype=np.float32) * f32(dt_prop)
    x = (pi * f0 * (t - t0)).astype(np.float32)
    a = (x * x).astype(np.float32)
    e = np.exp(-a.astype(np.float64)).astype(np.float32)
    S0 = ((f32(1.0) - f32(2.0) * a) * e).astype(np.float32)
    S0[t >= tmp] = 0.0
    return S0


def ricker_was(nt_prop: int, dt_prop: float, freq_ricker: float) -> np.ndarray:
    """Ricker_was (get_was_data.f90:466-495): the Ricker_mcs expression with the delay
    fixed at t0 = 1.5/f0 (FP32 quotient)."""
    return ricker_mcs(nt_prop, dt_prop, freq_ricker, np.float32(1.5) / np.float32(freq_ricker))


# ----------------------------------------------------------------------------------
# Device-resident form (torch tensors on the context's device)
# ----------------------------------------------------------------------------------

class DeviceShot:
    """Geometry of one shot for the device-resident form (psdm_rtm_geom)."""

    def __init__(self, ny, nx, nt, dPML, fs, dmodel, dt, nxSou, nySou, nxRec, nyRec, variant=0, store_snap=1,
                 nt_snap=None):
        self.arrays = [_i32(nxSou), _i32(nySou), _i32(nxRec), _i32(nyRec)]
        g = Geom()
        g.ny, g.nx, g.nt, g.dPML, g.fs, g.variant = ny, nx, nt, dPML, fs, variant
        g.store_snap = store_snap
        g.nt_snap = nt if nt_snap is None else nt_snap
        g.dmodel, g.dt = dmodel, dt
        g.NumSou, g.NumRec = self.arrays[0].size, self.arrays[2].size
        g.nxSou, g.nySou, g.nxRec, g.nyRec = (a.ctypes.data for a in self.arrays)
        self.g = g

    @property
    def shape(self):
        return self.g.ny, self.g.nx


def forward_device(ctx: Context, shot: DeviceShot, d_ctmp, d_S, d_rec, key: int):
    """Forward modelling on device: d_ctmp (nx, ny) torch tensor holding ctmp(ny,nx) in
    Fortran order, d_S (NumSou, nt) holding S(nt,NumSou), d_rec (NumRec, nt) receiving
    Rec_point(nt,NumRec).  Snapshots stay in HBM under ``key``."""
    rc = ctx._L.psdm_rtm_forward_device(ctx.h, C.byref(shot.g), C.c_void_p(d_ctmp.data_ptr()),
                                        C.c_void_p(d_S.data_ptr()), C.c_void_p(d_rec.data_ptr()), key)
    ctx.check(rc, "psdm_rtm_forward_device")


def backward_device(ctx: Context, shot: DeviceShot, d_ctmp, d_Sadj, d_kernel, key: int):
    """Back-propagation + imaging on device; d_kernel (nx, ny) accumulates kernel(ny,nx)."""
    rc = ctx._L.psdm_rtm_backward_device(ctx.h, C.byref(shot.g), C.c_void_p(d_ctmp.data_ptr()),
                                         C.c_void_p(d_Sadj.data_ptr()), C.c_void_p(d_kernel.data_ptr()), key)
    ctx.check(rc, "psdm_rtm_backward_device")
