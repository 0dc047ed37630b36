"""GPU parity of the Kirchhoff migration (psdm_rtm_eikonal, psdm_rtm_kirchhoff_image) with
the oracle (oracle/kirchhoff.c; its eikonal is pinned bit for bit to the reference's
solve_eikonal_FSM in tests/test_kirchhoff_oracle.py).

Bars: travel times bit for bit; the image within 1e-5 rel-L2 (the device's acosf / expf
differ from the host libm's by a few ulp; the travel-time index is exact)."""
import sys

import numpy as np
import pytest

from oracle import kirchhoff as K

pytestmark = pytest.mark.gpu
IMAGE_BAR = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def model(nx, ny, seed):
    rng = np.random.default_rng(seed)
    z = np.arange(ny)[None, :]
    v = np.where(z < max(1, ny // 5), 1500.0, 1700.0 + 2.0 * z * 10) * np.ones((nx, 1))
    v = v + 200.0 * rng.random((nx, ny))
    return v.astype(np.float32)


@pytest.mark.parametrize("nx,ny,srcs", [
    (40, 30, [(5, 1)]), (1, 5, [(1, 3)]), (7, 1, [(4, 1)]), (1, 1, [(1, 1)]), (64, 64, [(64, 64), (1, 1), (33, 2)]),
    (715, 601, [(300, 1)]),            # one MCS window, 768 threads x 1 column
    (1500, 40, [(700, 2), (1, 1)]),    # 2 columns per thread
    (3000, 12, [(2999, 12)]),          # 3 columns per thread (the 4-column kernel)
    (5000, 6, [(2500, 1)])])           # 5 columns per thread (the 8-column kernel)
def test_eikonal_bit_exact(ctx, nx, ny, srcs):
    import torch

    from psdm_amd.kirchhoff import travel_times

    say(f"[gpu] eikonal nx={nx} ny={ny} nsrc={len(srcs)}")
    v = model(nx, ny, nx + ny)
    T = travel_times(ctx, torch.from_numpy(v).cuda(), 10.0, [s[0] for s in srcs], [s[1] for s in srcs]).cpu().numpy()
    for q, (sx, sz) in enumerate(srcs):
        np.testing.assert_array_equal(T[q].view(np.uint32), K.eikonal(v, 10.0, 10.0, sx, sz).view(np.uint32))


def test_eikonal_shot_receivers_bit_exact(ctx):
    """A shot's receivers in one call (48 of them along the surface of a 715 x 601 window)."""
    import torch

    from psdm_amd.kirchhoff import travel_times

    nx, ny = 715, 601
    v = model(nx, ny, 3)
    xs = np.arange(40, 40 + 48 * 12, 12)
    zs = np.full(48, 2)
    say("[gpu] eikonal 48 receivers 715 x 601")
    T = travel_times(ctx, torch.from_numpy(v).cuda(), 10.0, xs, zs).cpu().numpy()
    for q in (0, 17, 47):
        np.testing.assert_array_equal(T[q].view(np.uint32), K.eikonal(v, 10.0, 10.0, int(xs[q]), 2).view(np.uint32))


def test_eikonal_batches_under_workspace_limit(ctx):
    """Many sources on a wide, shallow window with a workspace limit of three solves: the
    sources run in batches 3 + 3 + 3 + 2 that reuse one workspace, in the device form and in
    the host form (one batch of travel times on the device at a time); every field bit for
    bit the oracle's, both forms identical."""
    import ctypes as C

    import torch

    from psdm_amd.kirchhoff import eikonal_batch, set_eikonal_workspace, travel_times

    nx, ny, nsrc = 5102, 8, 11
    per = 4 * (2 * (nx + ny - 1) * nx + 1024)
    v = model(nx, ny, 21)
    rng = np.random.default_rng(4)
    xs = rng.integers(1, nx + 1, nsrc).astype(np.int32)
    xs[0], xs[-1] = 1, nx
    zs = rng.integers(1, ny + 1, nsrc).astype(np.int32)
    try:
        set_eikonal_workspace(ctx, 3 * per + per // 2)
        assert eikonal_batch(ctx, nx, ny, nsrc) == 3
        assert eikonal_batch(ctx, nx, ny, 2) == 2
        say(f"[gpu] eikonal batches of 3: nx={nx} ny={ny} nsrc={nsrc}")
        T = travel_times(ctx, torch.from_numpy(v).cuda(), 10.0, xs, zs).cpu().numpy()
        H = np.zeros((nsrc, nx, ny), np.float32)
        ctx.check(ctx._L.psdm_rtm_eikonal_host(ctx.h, v.ctypes.data_as(C.c_void_p), ny, nx, 10.0, 10.0, nsrc,
                                               xs.ctypes.data_as(C.c_void_p), zs.ctypes.data_as(C.c_void_p),
                                               H.ctypes.data_as(C.c_void_p)), "psdm_rtm_eikonal_host")
    finally:
        set_eikonal_workspace(ctx, 16 << 30)
    np.testing.assert_array_equal(H.view(np.uint32), T.view(np.uint32))
    for q in range(nsrc):
        np.testing.assert_array_equal(T[q].view(np.uint32), K.eikonal(v, 10.0, 10.0, int(xs[q]), int(zs[q])).view(np.uint32),
                                      err_msg=f"source {q}")


@pytest.mark.parametrize("dx,dy,odd", [(10.0, 10.0, False), (12.5, 7.0, False), (31.6821442, 31.6821442, False),
                                       (31.6821442, 5.0, False), (10.0, 10.0, True)])
def test_eikonal_constant_divisor_division_bit_exact(ctx, dx, dy, odd):
    """local_solver's quotients by dx*dx, dy*dy and 2*aa: a divisor that matched the IEEE
    division for every significand on the device uses a product and an FMA, any other keeps
    the IEEE division; either way the travel times are the oracle's bit for bit.  dx = 10
    (all three divisors pass the check) and f32(31.6821442)**2 = 1003.75824 (fails it: the
    two-operation form is off by one ulp for some significand) cover both paths.  `odd`:
    zero and near-zero velocities in the model (1/v infinite or huge: the numerators leave
    the range the check covers and take the IEEE division)."""
    import ctypes as C

    import torch

    from psdm_amd.kirchhoff import eikonal_division

    mask = eikonal_division(ctx, dx, dy)
    if dx == dy == 10.0:
        assert mask == 7
    if dx == 31.6821442:
        assert not mask & 1
    nx, ny = 150, 90
    v = model(nx, ny, 17)
    if odd:
        v[60:64, 30:33] = 0.0
        v[100, 10:20] = 1e-30
        v[20:25, 70] = 1e-19
    srcs = [(1, 1), (75, 2), (150, 90), (40, 61)]
    xs = np.array([q[0] for q in srcs], np.int32)
    zs = np.array([q[1] for q in srcs], np.int32)
    say(f"[gpu] eikonal dx={dx} dy={dy} divisor mask {mask}")
    d_v = torch.from_numpy(v).cuda()
    out = torch.empty((len(srcs), nx, ny), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.check(ctx._L.psdm_rtm_eikonal(ctx.h, C.c_void_p(d_v.data_ptr()), ny, nx, dx, dy, len(srcs),
                                      xs.ctypes.data_as(C.c_void_p), zs.ctypes.data_as(C.c_void_p),
                                      C.c_void_p(out.data_ptr())), "psdm_rtm_eikonal")
    T = out.cpu().numpy()
    for q, (sx, sz) in enumerate(srcs):
        np.testing.assert_array_equal(T[q].view(np.uint32), K.eikonal(v, dx, dy, sx, sz).view(np.uint32),
                                      err_msg=f"source {q}")


def test_eikonal_workspace_default_fits_a_wide_window(ctx):
    from psdm_amd.kirchhoff import eikonal_batch

    assert eikonal_batch(ctx, 715, 601, 481) == 481
    assert eikonal_batch(ctx, 5102, 801, 481) == 71


@pytest.mark.parametrize("nx,ny,nrec,nt,dt", [(120, 90, 24, 1500, 0.002), (301, 201, 60, 4000, 0.001)])
def test_image_vs_oracle(ctx, nx, ny, nrec, nt, dt):
    import torch

    from psdm_amd.kirchhoff import image, travel_times

    say(f"[gpu] kirchhoff image {nx}x{ny} nrec={nrec}")
    v = model(nx, ny, 7)
    rng = np.random.default_rng(nx)
    xr = np.linspace(5, nx - 5, nrec).astype(np.int32)
    dv = torch.from_numpy(v).cuda()
    tts = travel_times(ctx, dv, 10.0, [nx // 2], [1])[0]
    ttr = travel_times(ctx, dv, 10.0, xr, np.full(nrec, 2, np.int32))
    data = rng.standard_normal((nrec, nt)).astype(np.float32)
    got = image(ctx, tts, ttr, torch.from_numpy(data).cuda(), nt, 10.0, dt).cpu().numpy()
    want = K.image(tts.cpu().numpy(), ttr.cpu().numpy(), data, nt, 10.0, dt)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < IMAGE_BAR
    # the aperture and index logic agree exactly: the same pixels are zero
    np.testing.assert_array_equal(got == 0, want == 0)


def test_migrate_shot_and_errors(ctx):
    import torch

    from psdm_amd._native import RtmError
    from psdm_amd.kirchhoff import migrate_shot, travel_times

    say("[gpu] kirchhoff migrate_shot")
    nx, ny, nrec, nt, dt = 90, 70, 16, 1200, 0.002
    v = model(nx, ny, 11)
    data = np.random.default_rng(0).standard_normal((nrec, nt)).astype(np.float32)
    xr = np.arange(10, 10 + 4 * nrec, 4).astype(np.int32)
    zr = np.full(nrec, 1, np.int32)
    img = migrate_shot(ctx, torch.from_numpy(v).cuda(), 10.0, 45, 1, xr, zr, torch.from_numpy(data).cuda(), nt,
                       dt).cpu().numpy()
    tts = K.eikonal(v, 10.0, 10.0, 45, 1)
    ttr = np.stack([K.eikonal(v, 10.0, 10.0, int(x), 1) for x in xr])
    want = K.image(tts, ttr, data, nt, 10.0, dt)
    assert np.linalg.norm(img - want) / np.linalg.norm(want) < IMAGE_BAR
    with pytest.raises(RtmError, match="outside"):
        travel_times(ctx, torch.from_numpy(v).cuda(), 10.0, [nx + 1], [1])
    with pytest.raises(RtmError, match="8192"):
        travel_times(ctx, torch.zeros((8193, 2), device="cuda"), 10.0, [1], [1])
