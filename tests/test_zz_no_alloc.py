"""GPU test of psdm_rtm_psdm2twt_noalloc, the depth-to-time conversion whose workspace is
a static device array instead of a library allocation: through the C ABI, two calls in a
row (the second on a smaller model, so it runs on what the first left in the shared
scratch) give what psdm_rtm_psdm2twt gives, bit for bit, and the oracle's conversion;
a model beyond the scratch is refused before any device work."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import poststack as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def model(nz, nx, dd, seed):
    rng = np.random.default_rng(seed)
    z = np.arange(nz)[:, None] * dd
    wd = 400.0 + 50.0 * np.sin(np.arange(nx) / 7.0)
    vel = np.where(z < wd, 1500.0, 1800.0 + 1.2 * (z - wd)).astype(np.float32)
    vel[nz // 2:nz // 2 + 5, nx // 4] = 0.0  # a flat time run
    psdm = rng.standard_normal((nz, nx)).astype(np.float32)
    psdm[z[:, 0] < 300.0, :] = 0.0
    return vel, psdm


def convert(ctx, name, vel, psdm, dd, dt):
    """(nt, vel_twt, psdm_twt) from one entry point: a length query, then the full call."""
    import torch

    dv = torch.from_numpy(np.ascontiguousarray(vel.T)).to("cuda:0")
    dp = torch.from_numpy(np.ascontiguousarray(psdm.T)).to("cuda:0")
    nz, nx = vel.shape
    fn = getattr(ctx._L, name)
    nt = C.c_int()
    args = (ctx.h, nx, nz, dd, dt, C.c_void_p(dv.data_ptr()), C.c_void_p(dp.data_ptr()))
    ctx.check(fn(*args, 0, None, None, C.byref(nt)), name)
    vt = torch.full((nx, nt.value), float("nan"), device="cuda:0")
    pt = torch.full((nx, nt.value), float("nan"), device="cuda:0")
    nt2 = C.c_int()
    ctx.check(fn(*args, nt.value, C.c_void_p(vt.data_ptr()), C.c_void_p(pt.data_ptr()), C.byref(nt2)), name)
    torch.cuda.synchronize()
    assert nt2.value == nt.value
    return nt.value, vt.cpu().numpy().T, pt.cpu().numpy().T


def test_noalloc_equals_allocating_entry_point_twice(ctx):
    for nz, nx, dd, dt, seed in ((601, 300, 10.0, 0.004, 1), (157, 31, 12.5, 0.002, 2)):
        vel, psdm = model(nz, nx, dd, seed)
        say(f"[gpu] psdm2twt_noalloc {nz}x{nx} dd={dd} dt={dt}")
        n0, v0, p0 = convert(ctx, "psdm_rtm_psdm2twt_noalloc", vel, psdm, dd, dt)
        n1, v1, p1 = convert(ctx, "psdm_rtm_psdm2twt", vel, psdm, dd, dt)
        assert n0 == n1 == P.twt_length(P.twt_times(vel, dd), dt)
        np.testing.assert_array_equal(v0, v1)
        np.testing.assert_array_equal(p0, p1)
        with np.errstate(divide="ignore", invalid="ignore"):
            ov, op = P.psdm2twt(vel, psdm, dd, dt)
        np.testing.assert_array_equal(v0, ov)
        np.testing.assert_array_equal(p0, op)


def test_noalloc_refuses_a_model_beyond_its_scratch(ctx):
    from psdm_amd._native import RtmError

    nt = C.c_int()
    dummy = C.c_void_p(16)  # never dereferenced: the size check comes first
    with pytest.raises(RtmError, match="scratch"):
        ctx.check(ctx._L.psdm_rtm_psdm2twt_noalloc(ctx.h, 4096, 2049, 10.0, 0.004, dummy, dummy, 0, None, None,
                                                   C.byref(nt)), "psdm_rtm_psdm2twt_noalloc")
