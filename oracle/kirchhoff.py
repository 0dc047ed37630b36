"""Kirchhoff migration of one shot -- TEST INFRASTRUCTURE ONLY (the oracle).

ctypes wrappers of oracle/kirchhoff.c (the C restatement of solve_eikonal_FSM,
solve_eikonal_FSM.f90:1-83, and of the imaging loop of calculate_PSDM_kirchoff.f90:
394-447) and of the reference's own solve_eikonal_FSM compiled into oracle/_ref.
Arrays are (ny, nx) column-major, i.e. numpy (nx, ny) C-order views: one model column per
row.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import pyoracle as _po

F32 = np.float32
_P, _I, _F = C.c_void_p, C.c_int, C.c_float


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def eikonal(vel_cols: np.ndarray, dx: float, dy: float, sx: int, sz: int) -> np.ndarray:
    """vel_cols (nx, ny) (column i = row i); sx, sz 1-based -> T (nx, ny)."""
    L = _po.lib()
    fn = L.oracle_eikonal
    fn.argtypes = [_P, _I, _I, _F, _F, _I, _I, _P]
    fn.restype = None
    v = np.ascontiguousarray(vel_cols, F32)
    nx, ny = v.shape
    T = np.empty_like(v)
    fn(_p(v), ny, nx, dx, dy, sx, sz, _p(T))
    return T


def ref_eikonal(vel_cols: np.ndarray, dx: float, dy: float, sx: int, sz: int) -> np.ndarray:
    """The reference's solve_eikonal_FSM (oracle/_ref)."""
    L = _po.ref_lib()
    fn = L.ref_eikonal
    fn.argtypes = [_I, _I, _F, _F, _P, _I, _I, _P]
    fn.restype = None
    v = np.ascontiguousarray(vel_cols, F32)
    nx, ny = v.shape
    T = np.empty_like(v)
    fn(nx, ny, dx, dy, _p(v), sx, sz, _p(T))
    return T


def image(tts, ttr, data, nt_prop: int, dd: float, dt: float, vp0: float = 1500.0, theta_max: float = 60.0,
          theta_taper: float = 10.0, wmax: float = 1.0) -> np.ndarray:
    """tts (nx, ny), ttr (nrec, nx, ny), data (nrec, ld) -> image (nx, ny)."""
    L = _po.lib()
    fn = L.oracle_kirchhoff_image
    fn.argtypes = [_P, _P, _P, _I, _I, _I, _I, _I, _F, _F, _F, _F, _F, _F, _P]
    fn.restype = None
    tts = np.ascontiguousarray(tts, F32)
    ttr = np.ascontiguousarray(ttr, F32)
    data = np.ascontiguousarray(data, F32)
    nx, ny = tts.shape
    nrec, ld = data.shape
    out = np.empty_like(tts)
    fn(_p(tts), _p(ttr), _p(data), ld, nrec, nt_prop, ny, nx, dd, dt, vp0, theta_max, theta_taper, wmax, _p(out))
    return out
