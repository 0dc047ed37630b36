"""Depth-to-two-way-time conversion of the final image on the device (psdm_rtm_psdm2twt).

Mirrors psdm2twt (auxiliar_PSDM.f90:23-91) without its file writers: per model column the
two-way time of every depth sample, the uniform time axis up to the largest of them, and
the velocity and the image interpolated onto it by the reference's linear interpolation
(linear_interpolation.f90:72-103).  Device arrays are torch tensors: a model or image
``(nz, nx)`` in Fortran order is a contiguous ``(nx, nz)`` tensor (one column per row).
"""
from __future__ import annotations

import ctypes as C

from ._native import Context


def psdm2twt(ctx: Context, d_vel, d_psdm, dd: float, dt: float, out=None):
    """(vel_twt, psdm_twt) as (nx, nt) tensors from (nx, nz) d_vel, d_psdm (float32, same device).

    out: optional (vel_buf, psdm_buf), contiguous float32 tensors of nx * nt_cap elements
    each on the same device; then one call computes nt (which must not exceed nt_cap) and
    fills their first nx * nt elements, returned as (nx, nt) views.  Without it a length
    query runs first and the outputs are allocated to fit."""
    import torch

    for t in (d_vel, d_psdm):
        if t.dtype != torch.float32 or not t.is_contiguous() or t.dim() != 2:
            raise ValueError("d_vel / d_psdm must be contiguous (nx, nz) float32 tensors")
    if d_vel.shape != d_psdm.shape:
        raise ValueError("d_vel and d_psdm must have the same shape")
    nx, nz = d_vel.shape
    nt = C.c_int()
    args = (ctx.h, nx, nz, dd, dt, C.c_void_p(d_vel.data_ptr()), C.c_void_p(d_psdm.data_ptr()))
    if out is not None:
        vb, pb = out
        for t in (vb, pb):
            if t.dtype != torch.float32 or not t.is_contiguous() or t.device != d_vel.device:
                raise ValueError("out buffers must be contiguous float32 tensors on the inputs' device")
        cap = min(vb.numel(), pb.numel()) // nx
        ctx.check(ctx._L.psdm_rtm_psdm2twt(*args, cap, C.c_void_p(vb.data_ptr()), C.c_void_p(pb.data_ptr()),
                                           C.byref(nt)), "psdm_rtm_psdm2twt")
        n = nx * nt.value
        return vb.view(-1)[:n].view(nx, nt.value), pb.view(-1)[:n].view(nx, nt.value)
    ctx.check(ctx._L.psdm_rtm_psdm2twt(*args, 0, None, None, C.byref(nt)), "psdm_rtm_psdm2twt")
    vt = torch.empty((nx, nt.value), dtype=torch.float32, device=d_vel.device)
    pt = torch.empty_like(vt)
    ctx.check(ctx._L.psdm_rtm_psdm2twt(*args, nt.value, C.c_void_p(vt.data_ptr()), C.c_void_p(pt.data_ptr()),
                                       C.byref(nt)), "psdm_rtm_psdm2twt")
    return vt, pt


def timing(ctx: Context):
    """(time-accumulation ms, interpolation ms, conversions) per conversion since
    ctx.set_timing(n != 0)."""
    a, b, n = C.c_double(), C.c_double(), C.c_longlong()
    ctx.check(ctx._L.psdm_rtm_psdm2twt_timing(ctx.h, C.byref(a), C.byref(b), C.byref(n)), "psdm_rtm_psdm2twt_timing")
    return a.value, b.value, n.value
