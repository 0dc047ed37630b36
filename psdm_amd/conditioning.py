"""Gather conditioning before migration, on the device (psdm_rtm_condition).

Mirrors the per-shot steps of the reference driver between reading a gather and
migrating it (calculate_PSDM_RTM.f90:229-358): the Butterworth frequency filter
(freq_filter, auxiliar.f90:498-531), Akima resampling to the propagation step
(interpolation_akima.f90:1-31), the cut before the water-bottom reflection (cut_sf),
the cut before the first multiple (cut_multiple) and the half-cosine taper with zero
padding (taper_pad_data / taper_pad_trace, filters.f90:1-48).  The output is the
``Data_target(nt_taper, NumRec)`` gather the migration takes (in natural time order;
the engine and the one-shot solvers time-reverse it themselves).

Device arrays are torch tensors: a gather ``Field_Data(nt, NumRec)`` in Fortran order
is a contiguous ``(NumRec, nt)`` tensor (one trace per row).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import Context

_P, _I, _F = C.c_void_p, C.c_int, C.c_float


class Cond(C.Structure):
    """psdm_rtm_cond."""

    _fields_ = [
        ("nt_data", _I), ("dt_data", _F), ("typef", _I), ("f_init", _F), ("f_final", _F), ("nt_prop2", _I),
        ("dt_prop", _F), ("nt_taper", _I), ("taper_option", _I), ("N_taper", _I), ("N_zeros", _I),
        ("nt_sf", _P), ("nt_cut", _P),
    ]


def taper_lengths(T_taper: float, T_zeros: float, dt_prop: float):
    """N_taper = max(1, int(T_taper/dt_prop)), N_zeros = max(0, int(T_zeros/dt_prop)) in FP32
    (calculate_PSDM_RTM.f90:286-287)."""
    f = np.float32
    return max(1, int(f(f(T_taper) / f(dt_prop)))), max(0, int(f(f(T_zeros) / f(dt_prop))))


def cut_indices(nxSou: int, nySou: int, nxRec, nyRec, bathymetry, dmodel: float, dt_prop: float, kind: str):
    """Per-receiver sample index of the cut (calculate_PSDM_RTM.f90:229-282), FP32 as the
    reference: the symmetric water-bottom leg sqrt((offset/2)^2 + ((D_src + D_rec)/2)^2) at
    the midpoint's water depth; kind 'multiple' -> nt_cut = ceiling(0.9 * 4 leg / 1500 / dt),
    'sf' -> nt_sf = ceiling(1.1 * 2 leg / 1500 / dt).  Indices are 1-based model columns /
    rows; bathymetry(i) is the water depth (m) under model column i."""
    f = np.float32
    dm = f(dmodel)
    out = np.empty(len(nxRec), np.int32)
    for j in range(len(nxRec)):
        offset = abs(f(nxSou) * dm - f(nxRec[j]) * dm)
        mid_x = int(np.ceil(f(f(0.5) * f(nxSou + nxRec[j]))))
        bat = f(bathymetry[mid_x - 1])
        d_src = bat - f(nySou) * dm
        d_rec = bat - f(nyRec[j]) * dm
        leg = np.sqrt(f(f(offset * f(0.5)) ** 2 + f(f(0.5) * (d_src + d_rec)) ** 2))
        if kind == "multiple":
            out[j] = int(np.ceil(f(f(0.9) * (f(f(4.0) * leg) / f(1500.0))) / f(dt_prop)))
        elif kind == "sf":
            out[j] = int(np.ceil(f(f(1.1) * (f(f(2.0) * leg) / f(1500.0))) / f(dt_prop)))
        else:
            raise ValueError("kind must be 'multiple' or 'sf'")
    return out


def condition(ctx: Context, d_field, dt_data: float, nt_prop2: int, dt_prop: float, nt_taper: int | None = None,
              typef: int = 3, f_init: float = 0.0, f_final: float = 0.0, taper_option: int = 0, N_taper: int = 0,
              N_zeros: int = 0, nt_sf=None, nt_cut=None, out=None):
    """Data_target (NumRec, nt_taper) tensor from Field_Data d_field (NumRec, nt_data), both
    FP32 on ctx's device (d_field is not modified)."""
    import torch

    if d_field.dtype != torch.float32 or not d_field.is_contiguous() or d_field.dim() != 2:
        raise ValueError("d_field must be a contiguous (NumRec, nt_data) float32 tensor")
    nrec, nt_data = d_field.shape
    nt_taper = nt_prop2 if nt_taper is None else nt_taper
    if out is None:
        out = torch.empty((nrec, nt_taper), dtype=torch.float32, device=d_field.device)
    elif tuple(out.shape) != (nrec, nt_taper) or not out.is_contiguous() or out.dtype != torch.float32:
        raise ValueError("out must be a contiguous (NumRec, nt_taper) float32 tensor")
    sf = None if nt_sf is None else np.ascontiguousarray(nt_sf, dtype=np.int32)
    cut = None if nt_cut is None else np.ascontiguousarray(nt_cut, dtype=np.int32)
    for a in (sf, cut):
        if a is not None and a.size != nrec:
            raise ValueError("nt_sf / nt_cut need one entry per trace")
    q = Cond(nt_data, dt_data, typef, f_init, f_final, nt_prop2, dt_prop, nt_taper, taper_option, N_taper, N_zeros,
             sf.ctypes.data if sf is not None else None, cut.ctypes.data if cut is not None else None)
    ctx.check(ctx._L.psdm_rtm_condition(ctx.h, C.byref(q), nrec, C.c_void_p(d_field.data_ptr()),
                                        C.c_void_p(out.data_ptr())), "psdm_rtm_condition")
    return out


def timing(ctx: Context):
    """(filter ms, resample/cut/taper ms, calls) per call since ctx.set_timing(n != 0)."""
    f, r, n = C.c_double(), C.c_double(), C.c_longlong()
    ctx.check(ctx._L.psdm_rtm_condition_timing(ctx.h, C.byref(f), C.byref(r), C.byref(n)),
              "psdm_rtm_condition_timing")
    return f.value, r.value, n.value
