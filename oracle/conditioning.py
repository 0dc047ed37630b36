"""Gather conditioning before migration -- TEST INFRASTRUCTURE ONLY (the oracle).

CPU restatement of the per-shot data conditioning of the reference driver
(calculate_PSDM_RTM.f90:229-358), used by tests/ as the checker of the device path:

* ``butterworth_response`` -- the transfer function H(j), j = 1..ts/2+1, of the
  6th-order Butterworth filter built by filters_function1D (filters.f90:126-232):
  frequencies accumulated in FP32 exactly as the reference (f = f + fe/ts), the
  products of second-order sections evaluated in FP64;
* ``freq_filter`` -- per trace: zero pad to 24*nt with the trace at the end, forward FFT
  (normalised by 1/N, FFTPACK's cfft1f convention), multiply by H (mirrored to the
  negative frequencies as conj, filters.f90:227-229), inverse FFT, real part of the last
  nt samples (auxiliar.f90:498-531, filters.f90:94-111).  The FFT runs in FP64 (numpy)
  instead of FFTPACK's FP32 passes;
* ``akima`` -- interpolation_akima (interpolation_akima.f90:1-31) over ACM 433 as
  modified in akima433.f90:6-170, vectorised over the output samples in FP32 with the
  reference's operation order;
* ``taper_pad_trace`` -- the half-cosine taper and zero pad (filters.f90:20-48), with
  libm's cosf as the reference's cos;
* ``condition`` -- the driver's sequence: filter, resample (or copy), zero the first
  nt_sf samples, then taper/pad at nt_prop2 or cut at nt_cut (calculate_PSDM_RTM.f90:
  299-358);
* ``cut_indices`` -- nt_cut / nt_sf from the bathymetry (calculate_PSDM_RTM.f90:229-282).

Pinned by tests/test_conditioning_oracle.py against the reference's own routines
compiled into oracle/_ref and against closed forms.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import pyoracle as _po

ZERO_PAD = 24  # auxiliar.f90:507
F32 = np.float32

_libm = C.CDLL("libm.so.6")
_libm.cosf.argtypes = [C.c_float]
_libm.cosf.restype = C.c_float
_libm.atanf.argtypes = [C.c_float]
_libm.atanf.restype = C.c_float


def frequencies(ts: int, fe: float) -> np.ndarray:
    """f_j, j = 1..ts/2+1: f = 0, then f = f + fe/ts in FP32 (filters.f90:143-148)."""
    step = F32(F32(fe) / F32(ts))
    out = np.empty(ts // 2 + 1, F32)
    f = F32(0.0)
    for j in range(ts // 2 + 1):
        out[j] = f
        f = F32(f + step)
    return out


def butterworth_response(ts: int, fe: float, typef: int, f1: float, fc: float) -> np.ndarray:
    """H(j), j = 1..ts/2+1 (complex128), of filters_function1D with n = 6, parity 0,
    fs = 40 Hz for the high-pass, f2 = fc, f0 = sqrt(f1 f2), Q = f0/(f2-f1)
    (filters.f90:71-90, 126-217); H(1) = 1 for the low-pass, 0 otherwise (:221-225)."""
    n = 6
    pi = F32(3.14159265)
    fs = 40.0
    f0 = float(np.sqrt(F32(F32(f1) * F32(fc))))
    Q = float(F32(F32(f0) / F32(F32(fc) - F32(f1))))
    f = frequencies(ts, fe).astype(np.float64)
    B = np.ones(f.size, np.complex128)
    with np.errstate(divide="ignore", invalid="ignore"):
        for k in range(1, n // 2 + 1):
            c = float(np.cos(np.float64(F32(F32(2 * k + n - 1) * pi / F32(2 * n)))))
            if typef == 1:
                z = 1j * f / float(fc)
            elif typef == 2:
                z = fs / (1j * f)
            else:
                z = Q * ((1j * f) / f0 + f0 / (1j * f))
            B = B * (z ** 2 - 2 * z * c + 1)
        H = 1.0 / B
    H[0] = 1.0 if typef == 1 else 0.0
    return H


def filter_taps(nt: int, fe: float, typef: int, f1: float, fc: float) -> np.ndarray:
    """g[d], d = -(nt-1)..nt-1: the real part of the inverse DFT of the full H (so the
    filtered trace is sum_m g[n-m] x[m] over its own samples)."""
    ts = ZERO_PAD * nt
    Hh = butterworth_response(ts, fe, typef, f1, fc)
    H = np.empty(ts, np.complex128)
    H[:ts // 2 + 1] = Hh
    H[ts // 2 + 1:] = np.conj(Hh[1:ts // 2][::-1])
    g = np.fft.ifft(H).real
    d = np.arange(-(nt - 1), nt)
    return g[d % ts]


def freq_filter(data: np.ndarray, dt: float, typef: int, f1: float, f2: float) -> np.ndarray:
    """Filtered copy of data(nt, Num) (Fortran order), FFT in FP64 (auxiliar.f90:498-531)."""
    data = np.asarray(data, F32)
    nt, num = data.shape
    ts = ZERO_PAD * nt
    fe = F32(F32(1.0) / F32(dt))
    Hh = butterworth_response(ts, float(fe), typef, f1, f2)
    H = np.empty(ts, np.complex128)
    H[:ts // 2 + 1] = Hh
    H[ts // 2 + 1:] = np.conj(Hh[1:ts // 2][::-1])
    out = np.array(data, dtype=F32, order="F")
    for i in range(num):
        x = data[:, i]
        if not np.any(x != 0):  # is_not_cero (auxiliar.f90:113-130)
            continue
        pad = np.zeros(ts)
        pad[ts - nt:] = x
        y = np.fft.ifft(np.fft.fft(pad) * H).real  # fft/N then unnormalised inverse = numpy's pair
        out[:, i] = y[ts - nt:].astype(F32)
    return out


def akima(data: np.ndarray, dt_data: float, nt: int, dt: float) -> np.ndarray:
    """interpolation_akima: data(nt_data, NumRec) sampled at (i-1)*dt_data resampled at
    (i-1)*dt, i = 1..nt, with INTRPL (akima433.f90:6-170) in FP32."""
    Y_all = np.asarray(data, F32)
    L = Y_all.shape[0]
    X = np.arange(L).astype(F32) * F32(dt_data)
    U = np.arange(nt).astype(F32) * F32(dt)
    if L < 2:
        raise ValueError("INTRPL needs at least 2 input points")
    out = np.empty((nt, Y_all.shape[1]), F32, order="F")
    # 1-based interval index I: 1 below X(1), L+1 at or above X(L), else the first i
    # with U < X(i) (the binary search of :52-61); L == 2 always takes I = 2
    I = np.searchsorted(X, U, side="right") + 1 if L > 2 else np.full(nt, 2)
    J = np.clip(I, 2, L)
    Xp = lambda q: X[np.clip(q - 1, 0, L - 1)]  # X(q), q 1-based (clipped where unused)
    for r in range(Y_all.shape[1]):
        Yc = Y_all[:, r]
        Yp = lambda q: Yc[np.clip(q - 1, 0, L - 1)]
        X3, Y3, X4, Y4 = Xp(J - 1), Yp(J - 1), Xp(J), Yp(J)
        A3 = X4 - X3
        M3 = (Y4 - Y3) / A3
        z = np.zeros(nt, F32)
        if L == 2:
            M2 = M3.copy()
            M4 = M3.copy()
            A2 = z.copy()
            A4 = z.copy()
            X2 = z.copy()
            Y2 = z.copy()
            X5 = z.copy()
            Y5 = z.copy()
        else:
            X2, Y2 = Xp(J - 2), Yp(J - 2)
            A2 = np.where(J == 2, F32(0), X3 - X2)
            with np.errstate(divide="ignore", invalid="ignore"):
                M2 = np.where(J == 2, F32(0), (Y3 - Y2) / np.where(J == 2, F32(1), A2))
            X5, Y5 = Xp(J + 1), Yp(J + 1)
            A4 = np.where(J == L, F32(0), X5 - X4)
            with np.errstate(divide="ignore", invalid="ignore"):
                M4 = np.where(J == L, F32(0), (Y5 - Y4) / np.where(J == L, F32(1), A4))
            M2 = np.where(J == 2, (M3 + M3) - M4, M2)
            M4 = np.where(J == L, (M3 + M3) - M2, M4)
        with np.errstate(divide="ignore", invalid="ignore"):
            A1 = X2 - Xp(J - 3)
            M1 = np.where(J <= 3, (M2 + M2) - M3, (Y2 - Yp(J - 3)) / np.where(J <= 3, F32(1), A1))
            A5 = Xp(J + 2) - X5
            M5 = np.where(J >= L - 1, (M4 + M4) - M3, (Yp(J + 2) - Y5) / np.where(J >= L - 1, F32(1), A5))
        # slopes at the interval ends (:87-107)
        W2 = np.abs(M4 - M3)
        W3 = np.abs(M2 - M1)
        SW = W2 + W3
        zero = SW == 0
        W2 = np.where(zero, F32(0.5), W2)
        W3 = np.where(zero, F32(0.5), W3)
        SW = np.where(zero, F32(1.0), SW)
        T3 = (W2 * M2 + W3 * M3) / SW
        W3b = np.abs(M5 - M4)
        W4 = np.abs(M3 - M2)
        SWb = W3b + W4
        zero = SWb == 0
        W3b = np.where(zero, F32(0.5), W3b)
        W4 = np.where(zero, F32(0.5), W4)
        SWb = np.where(zero, F32(1.0), SWb)
        T4 = (W3b * M3 + W4 * M4) / SWb
        P0, Q0, Q1, A3e, M3e, T4e = X3.copy(), Y3.copy(), T3.copy(), A3.copy(), M3.copy(), T4.copy()
        hi = I == L + 1  # extrapolation above X(L) (:108-117)
        if np.any(hi):
            SA = A2 + A3
            t4 = F32(0.5) * ((M4 + M5) - ((A2 * (A2 - A3)) * (M2 - M3)) / (SA * SA))
            Q1 = np.where(hi, T4, Q1)
            T4e = np.where(hi, t4, T4e)
            P0 = np.where(hi, X4, P0)
            Q0 = np.where(hi, Y4, Q0)
            A3e = np.where(hi, A2, A3e)
            M3e = np.where(hi, M4, M3e)
        lo = I == 1  # extrapolation below X(1) (:118-125)
        if np.any(lo):
            SA = A3 + A4
            t3 = F32(0.5) * ((M1 + M2) - ((A4 * (A3 - A4)) * (M3 - M4)) / (SA * SA))
            T4e = np.where(lo, T3, T4e)
            Q1 = np.where(lo, t3, Q1)
            P0 = np.where(lo, X3 - A4, P0)
            Q0 = np.where(lo, Y3 - M2 * A4, Q0)
            A3e = np.where(lo, A4, A3e)
            M3e = np.where(lo, M2, M3e)
        Q2 = ((F32(2.0) * (M3e - Q1) + M3e) - T4e) / A3e
        Q3 = (((-M3e) - M3e + Q1) + T4e) / (A3e * A3e)
        DX = U - P0
        out[:, r] = Q0 + DX * (Q1 + DX * (Q2 + DX * Q3))
    return out


def taper_weights(N_taper: int) -> np.ndarray:
    """w(n) = 0.5*(1 + cos(pi*n/N_taper)), n = 0..N_taper-1, pi = 4 atan(1) (filters.f90:32-40)."""
    pi = F32(F32(4.0) * _libm.atanf(1.0))
    return np.array([F32(F32(0.5) * F32(F32(1.0) + _libm.cosf(float(F32(pi * F32(n)) / F32(N_taper)))))
                     for n in range(N_taper)], F32)


def taper_pad_trace(trace: np.ndarray, nt_prop: int, N_taper: int, N_zeros: int) -> np.ndarray:
    """filters.f90:20-48 on a copy (1-based nt_prop)."""
    t = np.array(trace, F32)
    w = taper_weights(N_taper)
    i0 = max(1, nt_prop - N_taper + 1)
    for n in range(N_taper):
        if i0 + n <= nt_prop:
            t[i0 + n - 1] = F32(t[i0 + n - 1] * w[n])
    if N_zeros > 0:
        t[nt_prop:] = 0
    return t


def condition(raw: np.ndarray, dt_data: float, typef: int, f_init: float, f_final: float, nt_prop2: int,
              dt_prop: float, nt_taper: int, N_taper: int = 0, N_zeros: int = 0, taper_option: int = 0,
              nt_sf=None, nt_cut=None) -> np.ndarray:
    """Data_target(nt_taper, NumRec) of calculate_PSDM_RTM.f90:299-358 from Field_Data."""
    raw = np.asarray(raw, F32)
    num = raw.shape[1]
    fd = freq_filter(raw, dt_data, typef, f_init, f_final) if (f_init > 0 and f_final > 0) else raw
    out = np.zeros((nt_taper, num), F32, order="F")
    if F32(dt_data) != F32(dt_prop):
        out[:nt_prop2] = akima(fd, dt_data, nt_prop2, dt_prop)
    else:
        out[:nt_prop2] = fd[:nt_prop2]
    if nt_sf is not None:
        for i in range(num):
            out[:nt_sf[i], i] = 0
    if nt_cut is None:
        if taper_option != 0:
            for i in range(num):
                out[:, i] = taper_pad_trace(out[:, i], nt_prop2, N_taper, N_zeros)
    else:
        for i in range(num):
            out[nt_cut[i] - 1:, i] = 0
            out[:, i] = taper_pad_trace(out[:, i], nt_cut[i], N_taper, N_zeros)
    return out


def cut_indices(nxSou: int, nySou: int, nxRec, nyRec, bathymetry, dmodel: float, dt_prop: float, kind: str):
    """nt_cut ('multiple', 0.9 x the first water-bottom multiple) or nt_sf ('sf', 1.1 x
    the water-bottom reflection) per receiver (calculate_PSDM_RTM.f90:229-282), FP32."""
    out = np.empty(len(nxRec), np.int32)
    dm = F32(dmodel)
    for j in range(len(nxRec)):
        offset = abs(F32(nxSou) * dm - F32(nxRec[j]) * dm)
        mid_x = int(np.ceil(F32(F32(0.5) * F32(nxSou + nxRec[j]))))
        D_src = F32(bathymetry[mid_x - 1]) - F32(nySou) * dm
        D_rec = F32(bathymetry[mid_x - 1]) - F32(nyRec[j]) * dm
        leg = np.sqrt(F32((offset * F32(0.5)) ** 2 + (F32(0.5) * (D_src + D_rec)) ** 2))
        if kind == "multiple":
            twt = F32(F32(4.0) * leg) / F32(1500.0)
            out[j] = int(np.ceil(F32(F32(0.9) * twt) / F32(dt_prop)))
        else:
            twt = F32(F32(2.0) * leg) / F32(1500.0)
            out[j] = int(np.ceil(F32(F32(1.1) * twt) / F32(dt_prop)))
    return out


# --- the reference's own routines (oracle/_ref) -------------------------------------

def _ref():
    L = _po.ref_lib()
    if not getattr(L, "_cond_bound", False):
        L.ref_freq_filter.argtypes = [C.c_int, C.c_float, C.c_int, C.c_int, C.c_float, C.c_float, C.c_void_p]
        L.ref_interpolation_akima.argtypes = [C.c_int, C.c_float, C.c_int, C.c_void_p, C.c_int, C.c_float,
                                              C.c_void_p]
        L.ref_taper_pad_trace.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        for f in (L.ref_freq_filter, L.ref_interpolation_akima, L.ref_taper_pad_trace):
            f.restype = None
        L._cond_bound = True
    return L


def ref_freq_filter(data, dt, typef, f1, f2):
    d = np.array(data, F32, order="F")
    nt, num = d.shape
    _ref().ref_freq_filter(nt, dt, num, typef, f1, f2, d.ctypes.data_as(C.c_void_p))
    return d


def ref_akima(data, dt_data, nt, dt):
    d = np.array(data, F32, order="F")
    out = np.zeros((nt, d.shape[1]), F32, order="F")
    _ref().ref_interpolation_akima(d.shape[0], dt_data, d.shape[1], d.ctypes.data_as(C.c_void_p), nt, dt,
                                   out.ctypes.data_as(C.c_void_p))
    return out


def ref_taper_pad_trace(trace, nt_prop, N_taper, N_zeros):
    t = np.array(trace, F32)
    _ref().ref_taper_pad_trace(t.ctypes.data_as(C.c_void_p), nt_prop, N_taper, N_zeros, t.size)
    return t
