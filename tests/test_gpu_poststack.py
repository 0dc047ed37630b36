"""GPU parity of the depth-to-time conversion (psdm_rtm_psdm2twt) with the oracle
(oracle/poststack.py, itself pinned to the reference's linear_interpolation_1D): bit for
bit, NaN included, on a water-over-gradient model and on columns with zero and negative
velocities."""
import sys

import numpy as np
import pytest

from oracle import poststack as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def device_twt(ctx, vel, psdm, dd, dt):
    import torch

    from psdm_amd.poststack import psdm2twt

    dv = torch.from_numpy(np.ascontiguousarray(vel.T)).to("cuda:0")
    dp = torch.from_numpy(np.ascontiguousarray(psdm.T)).to("cuda:0")
    vt, pt = psdm2twt(ctx, dv, dp, dd, dt)
    torch.cuda.synchronize()
    return vt.cpu().numpy().T, pt.cpu().numpy().T


def model(nz, nx, dd, seed):
    rng = np.random.default_rng(seed)
    z = np.arange(nz)[:, None] * dd
    wd = 400.0 + 50.0 * np.sin(np.arange(nx) / 7.0)
    vel = np.where(z < wd, 1500.0, 1800.0 + 1.2 * (z - wd)).astype(np.float32)
    psdm = rng.standard_normal((nz, nx)).astype(np.float32)
    psdm[z[:, 0] < 300.0, :] = 0.0  # no image in the shallow water
    psdm[:, nx // 3] = 0.0          # a column of zero fold
    return np.asfortranarray(vel), np.asfortranarray(psdm)


@pytest.mark.parametrize("nz,nx,dd,dt", [(300, 48, 10.0, 0.004), (157, 31, 12.5, 0.002)])
def test_psdm2twt_bit_exact_vs_oracle(ctx, nz, nx, dd, dt):
    vel, psdm = model(nz, nx, dd, nz)
    say(f"[gpu] psdm2twt {nz}x{nx} dd={dd} dt={dt}")
    gv, gp = device_twt(ctx, vel, psdm, dd, dt)
    ov, op = P.psdm2twt(vel, psdm, dd, dt)
    assert gv.shape == ov.shape
    np.testing.assert_array_equal(gv, ov)
    np.testing.assert_array_equal(gp, op)


def test_psdm2twt_zero_and_negative_velocities(ctx):
    """Zero velocities (flat time runs; image values there give the reference's 0*inf =
    NaN) and a negative velocity (a non-monotone time column, scanned as the reference)."""
    nz, nx, dd, dt = 120, 6, 10.0, 0.004
    vel, psdm = model(nz, nx, dd, 3)
    vel[40:50, 1] = 0.0
    vel[60, 2] = -1500.0
    vel[70:75, 3] = 0.0
    psdm[70:75, 3] = 0.0
    say("[gpu] psdm2twt zero / negative velocities")
    gv, gp = device_twt(ctx, vel, psdm, dd, dt)
    with np.errstate(divide="ignore", invalid="ignore"):
        ov, op = P.psdm2twt(vel, psdm, dd, dt)
    np.testing.assert_array_equal(gv, ov)
    np.testing.assert_array_equal(gp, op)


def test_psdm2twt_length_query_and_errors(ctx):
    import ctypes as C

    import torch

    from psdm_amd._native import RtmError

    vel, psdm = model(100, 8, 10.0, 1)
    dv = torch.from_numpy(np.ascontiguousarray(vel.T)).to("cuda:0")
    nt = C.c_int()
    args = (ctx.h, 8, 100, 10.0, 0.004, C.c_void_p(dv.data_ptr()), C.c_void_p(dv.data_ptr()))
    assert ctx._L.psdm_rtm_psdm2twt(*args, 0, None, None, C.byref(nt)) == 0
    assert nt.value == P.twt_length(P.twt_times(vel, 10.0), 0.004)
    out = torch.empty((8, 4), device="cuda:0")
    with pytest.raises(RtmError, match="nt_cap"):
        ctx.check(ctx._L.psdm_rtm_psdm2twt(*args, 4, C.c_void_p(out.data_ptr()), C.c_void_p(out.data_ptr()),
                                           C.byref(nt)), "psdm2twt")


def test_psdm2twt_single_call_into_given_buffers(ctx):
    """With output buffers of nt_cap >= nt rows one call fills the packed (nt, nx) result at
    their front: equal to the query-then-convert path, and to the oracle; a buffer too small
    for nt is rejected."""
    import torch

    from psdm_amd import poststack as PS
    from psdm_amd._native import RtmError

    nz, nx, dd, dt = 200, 23, 10.0, 0.004
    vel, psdm = model(nz, nx, dd, 5)
    dv = torch.from_numpy(np.ascontiguousarray(vel.T)).to("cuda:0")
    dp = torch.from_numpy(np.ascontiguousarray(psdm.T)).to("cuda:0")
    say("[gpu] psdm2twt into given buffers")
    vt, pt = PS.psdm2twt(ctx, dv, dp, dd, dt)
    nt = vt.shape[1]
    bufs = (torch.full((nx * (nt + 7),), -1.0, device="cuda:0"), torch.full((nx * (nt + 7),), -1.0, device="cuda:0"))
    v2, p2 = PS.psdm2twt(ctx, dv, dp, dd, dt, out=bufs)
    assert v2.shape == (nx, nt) and torch.equal(v2, vt) and torch.equal(p2, pt)
    ov, op = P.psdm2twt(vel, psdm, dd, dt)
    np.testing.assert_array_equal(v2.cpu().numpy().T, ov)
    np.testing.assert_array_equal(p2.cpu().numpy().T, op)
    small = (torch.empty(nx * (nt - 1), device="cuda:0"), torch.empty(nx * (nt - 1), device="cuda:0"))
    with pytest.raises(RtmError, match="nt_cap"):
        PS.psdm2twt(ctx, dv, dp, dd, dt, out=small)


@pytest.mark.parametrize("ny,nx,seed", [(601, 37, 0), (1, 1, 1), (255, 2001, 2)])
def test_normalize_bit_exact_vs_oracle(ctx, ny, nx, seed):
    """psdm_rtm_normalize (normalize_psdm, auxiliar_PSDM.f90:1-20) against the oracle's FP64
    normalize, bit for bit: folds 0 (image 0), 1 (the sum itself, rounded to FP32) and > 1,
    sums of every magnitude.  On FP32-representable sums the oracle equals the reference's
    own FP32 division (tests/test_poststack_oracle.py), so the device does too."""
    import ctypes as C

    import torch

    rng = np.random.default_rng(seed)
    s = rng.standard_normal((ny, nx)) * 10.0 ** rng.integers(-20, 20, (ny, nx))
    m32 = rng.random((ny, nx)) < 0.3
    s[m32] = s[m32].astype(np.float32)  # FP32-representable sums
    fold = rng.integers(0, 300, (ny, nx)).astype(np.int32)
    fold[rng.random((ny, nx)) < 0.25] = 0
    fold[rng.random((ny, nx)) < 0.25] = 1
    if ny * nx == 1:
        s[:], fold[:] = 1.0 / 3.0, 0
    say(f"[gpu] normalize {ny}x{nx} seed {seed}")
    ds = torch.from_numpy(np.ascontiguousarray(s.T)).to("cuda:0")
    df = torch.from_numpy(np.ascontiguousarray(fold.T)).to("cuda:0")
    out = torch.full((nx, ny), float("nan"), device="cuda:0")
    ctx.check(ctx._L.psdm_rtm_normalize(ctx.h, C.c_void_p(ds.data_ptr()), C.c_void_p(df.data_ptr()), ny, nx,
                                        C.c_void_p(out.data_ptr())), "psdm_rtm_normalize")
    torch.cuda.synchronize()
    got = out.cpu().numpy().T
    want = P.normalize(s, fold)
    np.testing.assert_array_equal(got, want)
    assert np.all(got[fold == 0] == 0)
    s32 = s.astype(np.float32)
    exact = s32.astype(np.float64) == s
    np.testing.assert_array_equal(got[exact], P.normalize_f32(s32, fold.astype(np.float32))[exact])
