"""Kirchhoff depth migration of one shot on the device (psdm_rtm_eikonal,
psdm_rtm_kirchhoff_image).

Mirrors calculate_PSDM_kirchoff (calculate_PSDM_kirchoff.f90:352-447): the window's
travel times from the source and from every receiver by fast sweeping (solve_eikonal_FSM,
via dynamic_window_tt_forward / _backward, dynamic_window.f90:230-325), then the
diffraction stack with the aperture-angle weight.  Device arrays are torch tensors; a
model ``(ny, nx)`` in Fortran order is a contiguous ``(nx, ny)`` tensor (one column per
row), travel times ``(nsrc, nx, ny)``, traces ``(nrec, nt)``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import Context


def travel_times(ctx: Context, d_vel, dd: float, src_x, src_z, out=None):
    """(nsrc, nx, ny) float32 travel times from the 1-based window points (src_x, src_z)."""
    import torch

    if d_vel.dtype != torch.float32 or not d_vel.is_contiguous() or d_vel.dim() != 2:
        raise ValueError("d_vel must be a contiguous (nx, ny) float32 tensor")
    torch.cuda.current_stream(d_vel.device).synchronize()
    nx, ny = d_vel.shape
    sx = np.ascontiguousarray(src_x, np.int32).ravel()
    sz = np.ascontiguousarray(src_z, np.int32).ravel()
    if sx.size != sz.size:
        raise ValueError("src_x and src_z must have the same length")
    if out is None:
        out = torch.empty((sx.size, nx, ny), dtype=torch.float32, device=d_vel.device)
    ctx.check(ctx._L.psdm_rtm_eikonal(ctx.h, C.c_void_p(d_vel.data_ptr()), ny, nx, dd, dd, sx.size,
                                      sx.ctypes.data_as(C.c_void_p), sz.ctypes.data_as(C.c_void_p),
                                      C.c_void_p(out.data_ptr())), "psdm_rtm_eikonal")
    return out


def set_eikonal_workspace(ctx: Context, max_bytes: int):
    """Limit of the eikonal workspace: the sources of travel_times run in batches that fit."""
    ctx.check(ctx._L.psdm_rtm_set_eikonal_workspace(ctx.h, int(max_bytes)), "psdm_rtm_set_eikonal_workspace")


def eikonal_batch(ctx: Context, nx: int, ny: int, nsrc: int) -> int:
    """Sources travel_times solves per launch for an (nx, ny) window under the limit."""
    nb = C.c_int(0)
    ctx.check(ctx._L.psdm_rtm_eikonal_batch(ctx.h, ny, nx, nsrc, C.byref(nb)), "psdm_rtm_eikonal_batch")
    return nb.value


def eikonal_division(ctx: Context, dx: float, dy: float) -> int:
    """Bit mask of the eikonal divisors (dx*dx, dy*dy, 2*aa) that use the two-operation
    division for these spacings (both forms are correctly rounded)."""
    m = C.c_int(0)
    ctx.check(ctx._L.psdm_rtm_eikonal_division(ctx.h, dx, dy, C.byref(m)), "psdm_rtm_eikonal_division")
    return m.value


def image(ctx: Context, d_tts, d_ttr, d_data, nt_prop: int, dd: float, dt: float, vp0: float = 1500.0,
          theta_max: float = 60.0, theta_taper: float = 10.0, wmax: float = 1.0, out=None):
    """(nx, ny) image of one shot: d_tts (nx, ny), d_ttr (nrec, nx, ny), d_data (nrec, ld)."""
    import torch

    for t in (d_tts, d_ttr, d_data):
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError("travel times and data must be contiguous float32 tensors")
    torch.cuda.current_stream(d_tts.device).synchronize()
    nx, ny = d_tts.shape
    nrec, ld = d_data.shape
    if tuple(d_ttr.shape) != (nrec, nx, ny):
        raise ValueError(f"d_ttr must be (nrec, nx, ny) = {(nrec, nx, ny)}")
    if out is None:
        out = torch.empty((nx, ny), dtype=torch.float32, device=d_tts.device)
    ctx.check(ctx._L.psdm_rtm_kirchhoff_image(ctx.h, C.c_void_p(d_tts.data_ptr()), C.c_void_p(d_ttr.data_ptr()),
                                              C.c_void_p(d_data.data_ptr()), ld, nrec, nt_prop, ny, nx, dd, dt, vp0,
                                              theta_max, theta_taper, wmax, C.c_void_p(out.data_ptr())),
              "psdm_rtm_kirchhoff_image")
    return out


def migrate_shot(ctx: Context, d_vel_w, dd: float, nxs: int, nys: int, nxr, nyr, d_data, nt_prop: int, dt: float,
                 vp0: float = 1500.0, theta_max: float = 60.0, theta_taper: float = 10.0, wmax: float = 1.0):
    """Image of one shot in its window: source (nxs, nys) and receivers (nxr, nyr), 1-based
    window indices (calculate_PSDM_kirchoff.f90:379-447)."""
    tts = travel_times(ctx, d_vel_w, dd, [nxs], [nys])[0]
    ttr = travel_times(ctx, d_vel_w, dd, nxr, nyr)
    return image(ctx, tts, ttr, d_data, nt_prop, dd, dt, vp0, theta_max, theta_taper, wmax)


def timing(ctx: Context):
    """(eikonal ms, imaging ms, eikonal calls, imaging calls) per call since ctx.set_timing(n != 0)."""
    a, b, na, nb = C.c_double(), C.c_double(), C.c_longlong(), C.c_longlong()
    ctx.check(ctx._L.psdm_rtm_kirchhoff_timing(ctx.h, C.byref(a), C.byref(b), C.byref(na), C.byref(nb)),
              "psdm_rtm_kirchhoff_timing")
    return a.value, b.value, na.value, nb.value
