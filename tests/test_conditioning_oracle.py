"""The gather-conditioning oracle (oracle/conditioning.py) pinned against the reference's
own routines compiled into oracle/_ref (freq_filter with FFTPACK, interpolation_akima /
INTRPL, taper_pad_trace), against the committed golden fixture, and against closed
forms; plus the host-side helpers of psdm_amd.conditioning."""
import os

import numpy as np
import pytest

from oracle import conditioning as O
from oracle import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "conditioning_ref.npz")
need_ref = pytest.mark.skipif(not pyoracle.ref_available(), reason="oracle/_ref not built")

AKIMA_CASES = [  # (L, dt_data, nt, dt): up- and down-sampling, outputs past X(L), tiny L
    (50, 0.004, 120, 0.0016), (300, 0.002, 700, 0.0008), (7, 0.01, 40, 0.003), (2, 0.01, 5, 0.004),
    (3, 0.01, 9, 0.004), (4, 0.01, 12, 0.004), (400, 0.0008, 150, 0.002), (60, 0.002, 60, 0.002)]


def _traces(L, n, seed):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((L, n)).astype(np.float32)
    d[:, 0] = np.sin(0.3 * np.arange(L))
    if n > 2:
        d[:, 1] = 0.0  # flat segments: the SW == 0 branches
        d[L // 2:, 1] = 1.0
    return np.asfortranarray(d)


@need_ref
@pytest.mark.parametrize("case", AKIMA_CASES)
def test_akima_bit_exact_vs_reference(case):
    L, dtd, nt, dt = case
    d = _traces(L, 3, L)
    np.testing.assert_array_equal(O.akima(d, dtd, nt, dt), O.ref_akima(d, dtd, nt, dt))


def test_akima_reproduces_linear_data():
    """Akima's slopes of a straight line are its slope, so the interpolant is the line
    itself up to the FP32 rounding of the samples (inside the data range; beyond X(L)
    the extrapolating cubic amplifies that rounding)."""
    L, dtd, nt, dt = 200, 0.002, 600, 0.0008
    x = np.arange(L) * dtd
    d = np.asfortranarray((3.0 - 2.5 * x).astype(np.float32)[:, None])
    v = O.akima(d, dtd, nt, dt)[:, 0]
    u = np.arange(nt) * np.float32(dt)
    inside = u < np.float32(L - 1) * np.float32(dtd)
    np.testing.assert_allclose(v[inside], (3.0 - 2.5 * u)[inside], rtol=0, atol=1e-6)
    np.testing.assert_allclose(v, 3.0 - 2.5 * u, rtol=0, atol=1e-3)


@need_ref
@pytest.mark.parametrize("nt_prop,N_taper,N_zeros,n", [(100, 20, 5, 130), (10, 20, 0, 30), (50, 1, 3, 60),
                                                         (64, 64, 0, 64)])
def test_taper_bit_exact_vs_reference(nt_prop, N_taper, N_zeros, n):
    t = np.random.default_rng(n).standard_normal(n).astype(np.float32)
    np.testing.assert_array_equal(O.taper_pad_trace(t, nt_prop, N_taper, N_zeros),
                                  O.ref_taper_pad_trace(t, nt_prop, N_taper, N_zeros))


def test_taper_weights_closed_form():
    w = O.taper_weights(16).astype(np.float64)
    np.testing.assert_allclose(w, 0.5 * (1 + np.cos(np.pi * np.arange(16) / 16)), rtol=0, atol=2e-7)


@need_ref
@pytest.mark.parametrize("typef,f1,f2", [(3, 5.0, 40.0), (1, 5.0, 30.0), (2, 5.0, 60.0)])
def test_filter_vs_reference_fftpack(typef, f1, f2):
    """FP64 FFT restatement vs the reference's FP32 FFTPACK evaluation."""
    d = _traces(400, 4, 7)
    d[:, 3] = 0.0  # an all-zero trace is left as it is (is_not_cero, auxiliar.f90:113-130)
    a = O.freq_filter(d, 0.002, typef, f1, f2)
    b = O.ref_freq_filter(d, 0.002, typef, f1, f2)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6
    assert not np.any(a[:, 3]) and not np.any(b[:, 3])


def test_butterworth_closed_forms():
    """|H|^2 = 1/(1 + (f/fc)^12) (low-pass), 1/(1 + (fs/f)^12) (high-pass, fs = 40 Hz),
    1/(1 + Q^12 (f/f0 - f0/f)^12) (band-pass), at the reference's FP32 frequencies; the
    reference's pole angles use pi rounded to FP32, hence the 1e-5 tolerance."""
    ts, fe = 4800, 500.0
    f = O.frequencies(ts, fe).astype(np.float64)
    H = O.butterworth_response(ts, fe, 1, 5.0, 30.0)
    np.testing.assert_allclose(np.abs(H) ** 2, 1 / (1 + (f / 30.0) ** 12), rtol=1e-5)
    H = O.butterworth_response(ts, fe, 2, 5.0, 60.0)
    np.testing.assert_allclose(np.abs(H[1:]) ** 2, 1 / (1 + (40.0 / f[1:]) ** 12), rtol=1e-5)
    assert H[0] == 0
    f1, f2 = 5.0, 40.0
    f0 = float(np.sqrt(np.float32(f1 * f2)))
    Q = float(np.float32(np.float32(f0) / np.float32(f2 - f1)))
    H = O.butterworth_response(ts, fe, 3, f1, f2)
    np.testing.assert_allclose(np.abs(H[1:]) ** 2, 1 / (1 + (Q * (f[1:] / f0 - f0 / f[1:])) ** 12), rtol=1e-5)
    # the FP32 frequency axis is the reference's running sum, not j*fe/ts
    assert O.frequencies(ts, fe)[-1] == np.float32(np.add.accumulate(np.full(ts // 2, np.float32(fe / ts)))[-1])


def test_filter_taps_equal_fft_filter():
    """The device's form of the filter, a convolution with taps g = Re(IDFT(H)) over the
    trace's own samples, equals the zero-padded FFT filter (both in FP64)."""
    nt, dt = 300, 0.002
    d = _traces(nt, 3, 3)
    fe = np.float32(np.float32(1.0) / np.float32(dt))
    g = O.filter_taps(nt, float(fe), 3, 5.0, 40.0)
    n = np.arange(nt)
    T = g[(n[:, None] - n[None, :]) + nt - 1]
    y = T @ d.astype(np.float64)
    ref = O.freq_filter(d, dt, 3, 5.0, 40.0)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 2e-7  # ref is rounded to FP32


@need_ref
def test_condition_pipeline_vs_reference_routines():
    """Driver sequence (filter, resample, cut_sf, cut_multiple + taper) against the same
    sequence over the reference's routines."""
    d = _traces(500, 5, 11)
    nt_sf = np.array([0, 30, 200, 5, 1200], np.int32)
    nt_cut = np.array([900, 1250, 400, 1000, 60], np.int32)
    kw = dict(dt_data=0.002, typef=3, f_init=5.0, f_final=40.0, nt_prop2=1200, dt_prop=0.0008, nt_taper=1250,
              N_taper=40, N_zeros=50)
    a = O.condition(d, nt_sf=nt_sf, nt_cut=nt_cut, **kw)
    f = O.ref_freq_filter(d, 0.002, 3, 5.0, 40.0)
    r = np.zeros((1250, 5), np.float32, order="F")
    r[:1200] = O.ref_akima(f, 0.002, 1200, 0.0008)
    for i in range(5):
        r[:nt_sf[i], i] = 0
        r[nt_cut[i] - 1:, i] = 0
        r[:, i] = O.ref_taper_pad_trace(r[:, i], nt_cut[i], 40, 50)
    assert np.linalg.norm(a - r) / np.linalg.norm(r) < 1e-6
    # no filter: bit for bit
    kw.update(f_init=0.0)
    a = O.condition(d, nt_sf=nt_sf, taper_option=1, **kw)
    r = np.zeros((1250, 5), np.float32, order="F")
    r[:1200] = O.ref_akima(d, 0.002, 1200, 0.0008)
    for i in range(5):
        r[:nt_sf[i], i] = 0
        r[:, i] = O.ref_taper_pad_trace(r[:, i], 1200, 40, 50)
    np.testing.assert_array_equal(a, r)


def test_golden_fixture_is_the_oracle():
    """The committed fixture (written from oracle/_ref) against the oracle restatement."""
    z = np.load(GOLDEN)
    for name in [k[:-4] for k in z.files if k.endswith("_raw")]:
        raw = np.asfortranarray(z[name + "_raw"])
        p = dict(zip(z["param_names"], z[name + "_params"]))
        nt_cut = z[name + "_nt_cut"] if name + "_nt_cut" in z.files else None
        nt_sf = z[name + "_nt_sf"] if name + "_nt_sf" in z.files else None
        out = O.condition(raw, float(p["dt_data"]), int(p["typef"]), float(p["f_init"]), float(p["f_final"]),
                          int(p["nt_prop2"]), float(p["dt_prop"]), int(p["nt_taper"]), int(p["N_taper"]),
                          int(p["N_zeros"]), int(p["taper_option"]), nt_sf=nt_sf, nt_cut=nt_cut)
        ref = z[name + "_out"]
        if float(p["f_init"]) > 0:
            assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-6, name
        else:
            np.testing.assert_array_equal(out, ref, err_msg=name)


def test_cut_indices_flat_bathymetry_closed_form():
    """Flat sea floor at depth D: the leg is sqrt((offset/2)^2 + (D - (zs+zr)/2)^2)."""
    from psdm_amd.conditioning import cut_indices

    dm, dt, D = 10.0, 0.0008, 1500.0
    nxSou, nySou = 400, 2
    nxRec = np.arange(420, 900, 25)
    nyRec = np.full(nxRec.size, 3)
    bat = np.full(2001, D, np.float32)
    for kind, fac, mult in (("multiple", 0.9, 4.0), ("sf", 1.1, 2.0)):
        got = cut_indices(nxSou, nySou, nxRec, nyRec, bat, dm, dt, kind)
        off = np.abs(nxSou - nxRec) * dm
        leg = np.sqrt((off / 2) ** 2 + (D - 0.5 * (nySou + nyRec) * dm) ** 2)
        want = np.ceil(fac * mult * leg / 1500.0 / dt).astype(np.int32)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(got, O.cut_indices(nxSou, nySou, nxRec, nyRec, bat, dm, dt, kind))


def test_taper_lengths():
    from psdm_amd.conditioning import taper_lengths

    assert taper_lengths(0.2, 0.1, 0.0008) == (250, 125)
    assert taper_lengths(0.0, 0.0, 0.001) == (1, 0)
