"""The Kirchhoff oracle (oracle/kirchhoff.c): the eikonal travel times pinned bit for bit
against the reference's solve_eikonal_FSM compiled into oracle/_ref; the imaging loop
(inline in the reference's driver, calculate_PSDM_kirchoff.f90:394-447, so not callable)
pinned by special cases that reduce it to sums written out independently here."""
import numpy as np
import pytest

from oracle import kirchhoff as K
from oracle import pyoracle

need_ref = pytest.mark.skipif(not pyoracle.ref_available(), reason="oracle/_ref not built")


@need_ref
@pytest.mark.parametrize("nx,ny,sx,sz,dx,dy,kind", [
    (40, 30, 5, 1, 10.0, 10.0, "random"), (61, 47, 61, 47, 10.0, 10.0, "random"), (23, 80, 12, 40, 12.5, 7.5, "random"),
    (1, 5, 1, 3, 10.0, 10.0, "random"), (7, 1, 4, 1, 10.0, 10.0, "random"), (1, 1, 1, 1, 5.0, 5.0, "random"),
    (120, 90, 60, 1, 10.0, 10.0, "layered"), (90, 120, 1, 120, 6.25, 6.25, "layered"),
    (50, 50, 25, 25, 10.0, 10.0, "lens"), (200, 61, 100, 3, 10.0, 10.0, "water")])
def test_eikonal_bit_exact_vs_reference(nx, ny, sx, sz, dx, dy, kind):
    rng = np.random.default_rng(nx * 1000 + ny)
    if kind == "random":
        v = 1500 + 3000 * rng.random((nx, ny))
    elif kind == "layered":
        v = np.tile(1500 + 25.0 * np.arange(ny), (nx, 1))
    elif kind == "lens":
        x, z = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        v = 3000 - 1200 * np.exp(-((x - 25) ** 2 + (z - 30) ** 2) / 60.0)
    else:  # water layer over a gradient, with a high-velocity block
        v = np.where(np.arange(ny)[None, :] < 20, 1500.0, 1800 + 30.0 * np.arange(ny)[None, :]) * np.ones((nx, 1))
        v[80:120, 30:45] = 4500.0
    v = v.astype(np.float32)
    a = K.eikonal(v, dx, dy, sx, sz)
    b = K.ref_eikonal(v, dx, dy, sx, sz)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    assert a[sx - 1, sz - 1] == 0.0 and (a > 0).sum() == nx * ny - 1


def test_eikonal_homogeneous_axis_distances():
    """Homogeneous medium, source at the surface: along the source column and the surface
    row every cell is reached from its axis neighbour, t = t_prev + dx/v (the quadratic has
    a = b = BIG or its one-sided root)."""
    nx, ny, v0, d = 30, 25, 2000.0, 10.0
    T = K.eikonal(np.full((nx, ny), v0, np.float32), d, d, 11, 1)
    step = np.float32(np.float32(1.0) / np.float32(v0)) * np.float32(d)
    col = T[10]
    for j in range(1, ny):
        assert abs(col[j] - col[j - 1] - step) <= 2e-7 * max(1.0, float(col[j]))
    row = T[:, 0]
    for i in range(11, nx):
        assert abs(row[i] - row[i - 1] - step) <= 2e-7 * max(1.0, float(row[i]))
    # off-axis cells are never faster than the straight ray
    x, z = np.meshgrid(np.arange(nx) - 10, np.arange(ny), indexing="ij")
    assert (T >= np.hypot(x, z) * d / v0 * (1 - 1e-6)).all()


def _case(rng, nx=12, ny=9, nrec=5, ld=400, dd=10.0, dt=0.002):
    tts = rng.uniform(0.0, 0.3, (nx, ny)).astype(np.float32)
    ttr = rng.uniform(0.0, 0.4, (nrec, nx, ny)).astype(np.float32)
    data = rng.standard_normal((nrec, ld)).astype(np.float32)
    return tts, ttr, data


def test_image_zero_data_and_out_of_range_index():
    rng = np.random.default_rng(1)
    tts, ttr, data = _case(rng)
    assert not K.image(tts, ttr, np.zeros_like(data), 400, 10.0, 0.002).any()
    # nt_prop = 2: every it = 1 + floor(t/dt) >= 2 is skipped
    assert not K.image(tts, ttr, data, 2, 10.0, 0.002).any()


def test_image_reduces_to_a_sum_of_central_differences():
    """theta_max = 90 (every angle passes, w_theta = 1) and a tiny wmax (the weight is
    min(cos/sqrt(rs*rr + eps), wmax) = wmax wherever cos > 0): the image is wmax times the
    sum over receivers of (d(it+1) - d(it-1))/(2 dt) with it = 1 + floor((ts + tr)/dt),
    skipping it <= 1, it >= nt_prop; row 0 (z = 0, cos = 0) is 0."""
    rng = np.random.default_rng(2)
    tts, ttr, data = _case(rng)
    nx, ny = tts.shape
    wmax, dt, nt_prop = np.float32(1e-12), np.float32(0.002), 300
    got = K.image(tts, ttr, data, nt_prop, 10.0, float(dt), theta_max=90.0, wmax=float(wmax))
    want = np.zeros((nx, ny), np.float32)
    for k in range(ttr.shape[0]):
        tot = (tts + ttr[k]).astype(np.float32)
        it = 1 + np.floor((tot / dt).astype(np.float32)).astype(np.int64)
        ok = (it > 1) & (it < nt_prop)
        itc = np.clip(it, 2, nt_prop - 1)
        amp = ((data[k][itc] - data[k][itc - 2]) / np.float32(2 * dt)).astype(np.float32)
        want = np.where(ok, (want + wmax * amp).astype(np.float32), want)
    want[:, 0] = 0.0
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-30)
    assert not got[:, 0].any()


def test_image_aperture_taper():
    """One receiver, one pixel per case, travel times chosen so that cos(theta) is known:
    the weight is w_theta * min(cos/sqrt(rs*rr + 1e-6), wmax) with w_theta = 1 inside
    theta_max and exp(-((theta - theta_max)/taper)^2) outside."""
    dd, dt, vp0 = 10.0, 0.001, 1500.0
    for theta_deg in (10.0, 59.0, 65.0, 80.0):
        z = 300.0  # pixel depth (row 30)
        # rs = rr = r with cos = z / (2 r)
        r = z / (2 * np.cos(np.radians(theta_deg)))
        t = np.float32(r / vp0)
        tts = np.full((1, 31), t, np.float32)
        ttr = np.full((1, 1, 31), t, np.float32)
        data = np.zeros((1, 2000), np.float32)
        it = 1 + int(np.floor(np.float32(np.float32(2 * t) / np.float32(dt))))
        data[0, it] = 1.0  # d(it+1) in the reference's 1-based indexing
        img = K.image(tts, ttr, data, 2000, dd, dt, vp0=vp0)
        rs = np.float32(vp0) * t
        c = float(np.float32(z) / np.float32(rs + rs + np.float32(1e-6)))
        th = np.degrees(np.arccos(c))
        w = 1.0 if th <= 60.0 else np.exp(-((th - 60.0) / 10.0) ** 2)
        want = w * min(c / np.sqrt(float(rs) * float(rs) + 1e-6), 1.0) / (2 * dt)
        assert abs(img[0, 30] - want) <= 1e-5 * abs(want), theta_deg
