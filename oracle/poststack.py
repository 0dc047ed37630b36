"""Depth-to-time conversion of the final image -- TEST INFRASTRUCTURE ONLY (the oracle).

CPU restatement of psdm2twt (auxiliar_PSDM.f90:23-91) without its file writers:

* ``twt_times`` -- per column the two-way time at each depth sample, accumulated in FP32
  in depth order: t(1) = 0, t(k) = t(k-1) + 2*dd/v(k) where |v(k)| > 0, else t(k-1)
  (:49-60);
* ``twt_length`` -- nt = 1 + ceiling(max |t| / dt) (:65-67);
* ``linear_interpolation_1d`` -- linear_interpolation.f90:72-103: constant extrapolation
  at both ends; inside, the first interval [x1(i), x1(i+1)] that contains the abscissa and
  whose two ordinates are both non-zero, y = y(i) + (a - x(i)) * ((y(i+1)-y(i))/(x(i+1)-x(i)))
  in FP32; 0 where no interval qualifies;
* ``psdm2twt`` -- velocity and image of every column on the time axis t(j) = (j-1)*dt;
* ``normalize_f32`` -- normalize_psdm (auxiliar_PSDM.f90:1-20) as the reference computes it:
  FP32 PSDM_sum / FP32 Cov_sum where Cov_sum > 0, else 0;
* ``normalize`` -- the same division on the FP64 stack (FP64 PSDM_sum, int32 fold) that the
  device keeps, rounded once to FP32.

Pinned by tests/test_poststack_oracle.py against the reference's own
linear_interpolation_1D and normalize_psdm compiled into oracle/_ref and against closed
forms.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import pyoracle as _po

F32 = np.float32


def twt_times(vel: np.ndarray, dd: float) -> np.ndarray:
    """time_total(nz, nx) of vel(nz, nx)."""
    vel = np.asarray(vel, F32)
    with np.errstate(divide="ignore"):
        inc = np.where(np.abs(vel) > 0, (F32(2.0) * F32(dd)) / vel, F32(0.0)).astype(F32)
    inc[0, :] = 0
    return np.add.accumulate(inc, axis=0, dtype=F32)  # sequential in depth order


def twt_length(t: np.ndarray, dt: float) -> int:
    return 1 + int(np.ceil(F32(F32(np.max(np.abs(t))) / F32(dt))))


def linear_interpolation_1d(x1, y1, x2) -> np.ndarray:
    x1 = np.asarray(x1, F32)
    y1 = np.asarray(y1, F32)
    x2 = np.asarray(x2, F32)
    n1 = x1.size
    y2 = np.zeros(x2.size, F32)
    for j, a3 in enumerate(x2):
        if a3 <= x1[0]:
            y2[j] = y1[0]
        elif a3 >= x1[n1 - 1]:
            y2[j] = y1[n1 - 1]
        else:
            for i in range(n1 - 1):
                a1, a2, b1, b2 = x1[i + 1], x1[i], y1[i + 1], y1[i]
                if a3 >= a2 and a3 <= a1 and abs(b1) > 0 and abs(b2) > 0:
                    m = F32(F32(b1 - b2) / F32(a1 - a2))
                    y2[j] = F32(b2 + F32(F32(a3 - a2) * m))
                    break
    return y2


def psdm2twt(vel: np.ndarray, psdm: np.ndarray, dd: float, dt: float):
    """(vel_twt, psdm_twt), both (nt, nx) FP32."""
    t = twt_times(vel, dd)
    nt = twt_length(t, dt)
    time = np.arange(nt).astype(F32) * F32(dt)
    nx = vel.shape[1]
    vt = np.empty((nt, nx), F32, order="F")
    pt = np.empty((nt, nx), F32, order="F")
    for i in range(nx):
        vt[:, i] = linear_interpolation_1d(t[:, i], vel[:, i], time)
        pt[:, i] = linear_interpolation_1d(t[:, i], psdm[:, i], time)
    return vt, pt


def ref_linear_interpolation_1d(x1, y1, x2) -> np.ndarray:
    """The reference's linear_interpolation_1D (oracle/_ref)."""
    L = _po.ref_lib()
    if not getattr(L, "_lin_bound", False):
        L.ref_linear_interpolation_1d.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        L.ref_linear_interpolation_1d.restype = None
        L._lin_bound = True
    x1 = np.ascontiguousarray(x1, F32)
    y1 = np.ascontiguousarray(y1, F32)
    x2 = np.ascontiguousarray(x2, F32)
    y2 = np.zeros(x2.size, F32)
    L.ref_linear_interpolation_1d(x1.ctypes.data_as(C.c_void_p), y1.ctypes.data_as(C.c_void_p), x1.size,
                                  x2.ctypes.data_as(C.c_void_p), y2.ctypes.data_as(C.c_void_p), x2.size)
    return y2


def normalize_f32(psdm_sum, cov_sum) -> np.ndarray:
    """normalize_psdm with FP32 PSDM_sum and FP32 Cov_sum (auxiliar_PSDM.f90:9-18)."""
    s = np.asarray(psdm_sum, F32)
    c = np.asarray(cov_sum, F32)
    out = np.zeros(s.shape, F32)
    pos = c > 0
    out[pos] = s[pos] / c[pos]
    return out


def normalize(psdm_sum, fold) -> np.ndarray:
    """PSDM = float32(PSDM_sum / fold) where fold > 0, else 0, from the FP64 sum and the
    int32 fold.  For an FP32-representable sum this equals normalize_f32: FP64 division
    followed by one rounding to FP32 is the correctly rounded FP32 quotient (53 >= 2*24+2
    significand bits, so the double rounding is innocuous)."""
    s = np.asarray(psdm_sum, np.float64)
    f = np.asarray(fold, np.int32)
    out = np.zeros(s.shape, F32)
    pos = f > 0
    out[pos] = (s[pos] / f[pos].astype(np.float64)).astype(F32)
    return out


def ref_normalize_psdm(psdm_sum, cov_sum) -> np.ndarray:
    """The reference's normalize_psdm (oracle/_ref) on (ny, nx) FP32 arrays."""
    L = _po.ref_lib()
    if not getattr(L, "_norm_bound", False):
        L.ref_normalize_psdm.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ref_normalize_psdm.restype = None
        L._norm_bound = True
    s = np.asfortranarray(psdm_sum, F32)
    c = np.asfortranarray(cov_sum, F32)
    out = np.zeros(s.shape, F32, order="F")
    ny, nx = s.shape
    L.ref_normalize_psdm(nx, ny, s.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p),
                         out.ctypes.data_as(C.c_void_p))
    return out
