"""Pins the oracle (oracle/rtm_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible slip (a dropped term, a wrong sign or index, a
transposed table) fails it: closed forms from the literature, brute force on tiny
grids, exact invariants, and the golden vectors written by the reference Fortran itself
(tests/golden/make_golden.py).
"""
import os
import sys
from fractions import Fraction

import numpy as np
import pytest

from oracle import pyoracle as O
from tests.cases import CASES, case, rel_l2

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _fornberg_central(deriv, radius):
    """Central-difference weights w_k (k = -r..r) exact for polynomials up to degree 2r,
    by solving the Taylor moment system sum_k w_k k^m = m! delta(m, deriv) in rationals."""
    ks = list(range(-radius, radius + 1))
    n = len(ks)
    A = [[Fraction(k) ** m for k in ks] for m in range(n)]
    b = [Fraction(0)] * n
    fact = 1
    for m in range(1, deriv + 1):
        fact *= m
    b[deriv] = Fraction(fact)
    # Gauss-Jordan in exact arithmetic
    M = [row + [rhs] for row, rhs in zip(A, b)]
    for col in range(n):
        piv = next(r for r in range(col, n) if M[r][col] != 0)
        M[col], M[piv] = M[piv], M[col]
        pv = M[col][col]
        M[col] = [x / pv for x in M[col]]
        for r in range(n):
            if r != col and M[r][col] != 0:
                f = M[r][col]
                M[r] = [x - f * y for x, y in zip(M[r], M[col])]
    return {k: M[i][n] for i, k in enumerate(ks)}


def test_fd_weights_are_fornberg_radius3():
    w1 = _fornberg_central(1, 3)
    w2 = _fornberg_central(2, 3)
    c1 = O.fd_coef(1)
    c2 = O.fd_coef(2)
    # first derivative: D = sum_k c1[k] (p(i+k) - p(i-k))  ->  c1[k] = w1[k]
    for k in range(1, 4):
        assert c1[k] == pytest.approx(float(w1[k]), rel=1e-15)
    assert c1[0] == 0.0
    # second derivative: D = sum_k c2[k] (p(i+k) + p(i-k)) with c2[0] the halved centre weight
    assert 2 * c2[0] == pytest.approx(float(w2[0]), rel=1e-15)
    for k in range(1, 4):
        assert c2[k] == pytest.approx(float(w2[k]), rel=1e-15)


def test_fd_weights_match_reference():
    t = np.load(os.path.join(GOLD, "tables.npz"))
    np.testing.assert_array_equal(O.fd_coef(1, "f32"), t["fd1"])
    np.testing.assert_array_equal(O.fd_coef(2, "f32"), t["fd2"])


def _cpml_literature(dPML, dt, d):
    """C-PML profile of Komatitsch & Martin (2007, Geophysics 72(5) SM155): damping
    d(x) = d0 (x/L)^N, d0 = -(N+1) vp ln(R) / (2L); frequency shift a(x) = a_max (1 - x/L);
    b = exp(-(d + a) dt), coefficient = d (b - 1) / (d + a).  The reference picks
    N = 2, vp = 2000 m/s, R = 1e-5, a_max = 30 pi, L = (dPML - 1) d, x = 0 at the first
    damped cell (solver_PML_coeffs.f90:16-29)."""
    N, vp, R, amax = 2, 2000.0, 1e-5, 30 * 3.14159265
    L = (dPML - 1) * d
    x = np.arange(dPML) * d
    d0 = -(N + 1) * vp * np.log(R) / (2 * L)
    damp = d0 * (x / L) ** N
    alpha = amax * (1 - x / L)
    b = np.exp(-(damp + alpha) * dt)
    a = damp * (b - 1) / (damp + alpha)
    return b, a


@pytest.mark.parametrize("dPML,dt,d", [(10, 0.001, 10.0), (40, 0.0008, 10.0), (40, 0.001, 12.5)])
def test_pml_profile_closed_form_and_reference(dPML, dt, d):
    fA, fB, cA, cB = O.pml_tables(dPML, dt, d, mode=0, precision="f64")
    dn = dPML + 6
    # the solver receives dt and dmodel as FP32 values
    b, a = _cpml_literature(dPML, float(np.float32(dt)), float(np.float32(d)))
    # first acc = 3 cells are undamped (A = B = 0), then the profile, then held constant
    assert np.all(fA[:3] == 0) and np.all(fB[:3] == 0)
    np.testing.assert_allclose(fA[3:3 + dPML], b, rtol=1e-12)
    np.testing.assert_allclose(fB[3:3 + dPML], a, rtol=1e-12, atol=1e-300)
    assert np.all(fA[3 + dPML:] == fA[2 + dPML]) and np.all(fB[3 + dPML:] == fB[2 + dPML])
    assert len(fA) == dn
    # against the reference's own PML_coeffs (FP32): the FP32 mode is bit-identical,
    # the FP64 mode (what the GPU library uses) rounds to within a few ulp
    t = np.load(os.path.join(GOLD, "tables.npz"))
    key = f"d{dPML}_{dt}_{d}"
    gA, gB, gcA, gcB = O.pml_tables(dPML, dt, d, mode=1, precision="f32")
    np.testing.assert_array_equal(gA, t[key + "_fA"])
    np.testing.assert_array_equal(gB, t[key + "_fB"])
    np.testing.assert_array_equal(gcA, t[key + "_cA"])
    np.testing.assert_array_equal(gcB, t[key + "_cB"])
    np.testing.assert_allclose(fA.astype(np.float32), t[key + "_fA"], rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(fB.astype(np.float32), t[key + "_fB"], rtol=1e-5, atol=1e-9)


def test_corner_tables_are_max_distance_rule():
    """PML_coeffs builds the corners with nested loops; they equal F(max(dist_x, dist_y)),
    dist = 1 next to the interior (the rule the GPU kernel evaluates)."""
    dPML = 12
    dn = dPML + 6
    fA, fB, cA, cB = O.pml_tables(dPML, 0.001, 10.0, mode=0, precision="f64")
    j = np.arange(dn)[:, None]
    i = np.arange(dn)[None, :]
    d_in = lambda v: v + 1          # yd rows / xr cols: index 0 is next to the interior
    d_out = lambda v: dn - v        # yu rows / xl cols: index dn-1 is next to the interior
    rules = {0: (d_in(j), d_out(i)),   # ld: bottom rows, left cols
             1: (d_in(j), d_in(i)),    # rd
             2: (d_out(j), d_out(i)),  # lu
             3: (d_out(j), d_in(i))}   # ru
    for q, (dy, dx) in rules.items():
        m = np.maximum(dy, dx) - 1
        np.testing.assert_array_equal(cA[q], fA[m])
        np.testing.assert_array_equal(cB[q], fB[m])


def test_derivative_passes_brute_force_and_reference():
    t = np.load(os.path.join(GOLD, "tables.npz"))
    p = t["deriv_in"].astype(np.float64)
    ny, nx = p.shape
    c1 = O.fd_coef(1)
    c2 = O.fd_coef(2)
    # brute force: interior points by the stencil sum, then the clamp copies
    bf = np.zeros((4, ny, nx))
    for j in range(ny):
        for i in range(3, nx - 3):
            bf[0, j, i] = sum(c2[k] * (p[j, i + k] + p[j, i - k]) for k in range(4))
            bf[2, j, i] = sum(c1[k] * (p[j, i + k] - p[j, i - k]) for k in range(4))
    for j in range(3, ny - 3):
        for i in range(nx):
            bf[1, j, i] = sum(c2[k] * (p[j + k, i] + p[j - k, i]) for k in range(4))
            bf[3, j, i] = sum(c1[k] * (p[j + k, i] - p[j - k, i]) for k in range(4))
    for w in (0, 2):
        bf[w, :, :3] = bf[w, :, 3:4]
        bf[w, :, nx - 3:] = bf[w, :, nx - 4:nx - 3]
    for w in (1, 3):
        bf[w, :3, :] = bf[w, 3:4, :]
        bf[w, ny - 3:, :] = bf[w, ny - 4:ny - 3, :]
    for w in range(4):
        d = O.deriv(w, p)
        np.testing.assert_allclose(d, bf[w], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(d, t[f"deriv_{w}"], rtol=1e-5, atol=1e-5)


def test_derivatives_exact_on_polynomials():
    """Radius-3 central differences are exact for polynomials up to degree 6."""
    ny, nx = 20, 24
    y = np.arange(ny, dtype=np.float64)[:, None] * np.ones((1, nx))
    x = np.ones((ny, 1)) * np.arange(nx, dtype=np.float64)[None, :]
    p = 0.01 * x ** 5 - 0.3 * x ** 2 + 2 * y ** 4 + x * y
    d2x = O.deriv(0, p)
    d1y = O.deriv(3, p)
    exact_d2x = 0.2 * x ** 3 - 0.6
    exact_d1y = 8 * y ** 3 + x
    np.testing.assert_allclose(d2x[:, 3:nx - 3], exact_d2x[:, 3:nx - 3], rtol=1e-9, atol=1e-7)
    np.testing.assert_allclose(d1y[3:ny - 3, :], exact_d1y[3:ny - 3, :], rtol=1e-9, atol=1e-7)


def test_zero_sum_weights_brute_force_and_constant_field():
    """The zero-sum form of the second-derivative passes (which 4 / 5): brute force of
    sum_k ((p(i+k) - p(i)) + (p(i-k) - p(i))) * w_k with the same clamp copies, the same
    operator as the reference form in exact arithmetic (FP64 agreement to rounding), and
    an exact zero on a constant field in FP32, where the reference form's FP32 weights
    leave -5.9e-8 per unit value (their sum is not zero)."""
    t = np.load(os.path.join(GOLD, "tables.npz"))
    p = t["deriv_in"].astype(np.float64)
    ny, nx = p.shape
    c2 = O.fd_coef(2)
    bf = np.zeros((2, ny, nx))
    for j in range(ny):
        for i in range(3, nx - 3):
            bf[0, j, i] = sum(c2[k] * ((p[j, i + k] - p[j, i]) + (p[j, i - k] - p[j, i])) for k in range(1, 4))
    for j in range(3, ny - 3):
        for i in range(nx):
            bf[1, j, i] = sum(c2[k] * ((p[j + k, i] - p[j, i]) + (p[j - k, i] - p[j, i])) for k in range(1, 4))
    bf[0, :, :3] = bf[0, :, 3:4]
    bf[0, :, nx - 3:] = bf[0, :, nx - 4:nx - 3]
    bf[1, :3, :] = bf[1, 3:4, :]
    bf[1, ny - 3:, :] = bf[1, ny - 4:ny - 3, :]
    for w in (0, 1):
        dz = O.deriv(4 + w, p)
        np.testing.assert_allclose(dz, bf[w], rtol=1e-12, atol=1e-12)
        # same operator as the reference form (centre weight -2 (w1 + w2 + w3) = 2 w0)
        np.testing.assert_allclose(dz, O.deriv(w, p), rtol=1e-12, atol=1e-12 * np.abs(p).max())
    const = np.full((ny, nx), 1.7, np.float32)
    for w in (4, 5):
        assert not O.deriv(w, const, precision="f32").any()
    assert np.abs(O.deriv(0, const, precision="f32")).max() > 0  # the reference form's residual
    w32 = O.fd_coef(2, precision="f32").astype(np.float64)
    assert abs(w32[0] + w32[1] + w32[2] + w32[3]) > 1e-8  # its weights do not sum to zero


def test_zero_sum_weights_exact_on_polynomials():
    ny, nx = 20, 24
    y = np.arange(ny, dtype=np.float64)[:, None] * np.ones((1, nx))
    x = np.ones((ny, 1)) * np.arange(nx, dtype=np.float64)[None, :]
    p = 0.01 * x ** 6 - 0.3 * x ** 2 + 2 * y ** 4 + x * y + 5.0
    np.testing.assert_allclose(O.deriv(4, p)[:, 3:nx - 3], (0.3 * x ** 4 - 0.6)[:, 3:nx - 3], rtol=1e-9, atol=1e-7)
    np.testing.assert_allclose(O.deriv(5, p)[3:ny - 3, :], (24 * y ** 2)[3:ny - 3, :], rtol=1e-9, atol=1e-7)


@pytest.mark.parametrize("name", ["ram_fs0", "disk_fs1", "ram_ss2"])
def test_zero_sum_solver_fp64_equals_reference_form_fp64(name):
    """In FP64 the zero-sum solver (variant bit 4) and the reference form agree to
    rounding; in FP32 they differ (the zero-sum form is not the reference's arithmetic)."""
    c = case(name)
    geo = (c["nxSou"], c["nySou"], c["nxRec"], c["nyRec"])
    out = {}
    for prec in ("f64", "f32"):
        for zs in (0, 4):
            v = c["variant"] | zs
            Rec, F = O.forward(*geo, c["S"], c["dx"], c["dt"], c["ctmp"], c["store_snap"], c["nt_snap"], c["nt"],
                               c["dPML"], c["fs"], variant=v, table_mode=0, precision=prec)
            k = O.backward(*geo, np.asfortranarray(Rec[::-1, :].astype(np.float32)), c["dx"], c["dt"], c["ctmp"],
                           c["store_snap"], c["nt_snap"], c["nt"], c["dPML"], c["fs"], Fields=F, variant=v,
                           table_mode=0, precision=prec)
            out[prec, zs] = (Rec, k)
    for q in (0, 1):
        assert rel_l2(out["f64", 4][q], out["f64", 0][q]) <= 1e-12
        assert rel_l2(out["f32", 4][q], out["f64", 0][q]) <= 1e-5
    assert not np.array_equal(out["f32", 4][0], out["f32", 0][0])


def _run(c, impl="oracle", table_mode=1, precision="f32", scale=1.0):
    geo = (c["nxSou"], c["nySou"], c["nxRec"], c["nyRec"])
    Rec, Fields = O.forward(*geo, c["S"] * scale, c["dx"], c["dt"], c["ctmp"], c["store_snap"], c["nt_snap"], c["nt"],
                            c["dPML"], c["fs"], variant=c["variant"], table_mode=table_mode, precision=precision,
                            impl=impl)
    Sadj = np.asfortranarray(Rec[::-1, :].astype(np.float32))
    kernel = O.backward(*geo, Sadj, c["dx"], c["dt"], c["ctmp"], c["store_snap"], c["nt_snap"], c["nt"], c["dPML"],
                        c["fs"], Fields=Fields, variant=c["variant"], table_mode=table_mode, precision=precision,
                        impl=impl)
    return Rec, Fields, kernel


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_bit_exact_vs_reference_golden(name):
    g = np.load(os.path.join(GOLD, f"{name}.npz"))
    c = case(name)
    # the stored inputs are the ones the case generator makes
    np.testing.assert_array_equal(g["ctmp"], c["ctmp"])
    Rec, Fields, kernel = _run(c, table_mode=1)
    np.testing.assert_array_equal(Rec, g["Rec"])
    np.testing.assert_array_equal(kernel, g["kernel"])
    if "snaps" in g:
        np.testing.assert_array_equal(Fields[:, :, g["snap_idx"]], g["snaps"])


@pytest.mark.parametrize("name", ["ram_fs0", "ram_fs1", "disk_fs1", "ram_ss2"])
def test_oracle_fp64_tables_and_fp64_twin_close(name):
    g = np.load(os.path.join(GOLD, f"{name}.npz"))
    c = case(name)
    Rec0, _, k0 = _run(c, table_mode=0)  # FP64 coefficient setup, FP32 stepping (the GPU's arithmetic)
    assert rel_l2(Rec0, g["Rec"]) < 1e-5
    assert rel_l2(k0, g["kernel"]) < 1e-5
    R64, _, k64 = _run(c, table_mode=0, precision="f64")
    assert rel_l2(g["Rec"], R64) < 1e-4
    assert rel_l2(g["kernel"], k64) < 1e-4


def test_oracle_linearity_exact():
    """Scaling the source by 2 scales every intermediate by 2 exactly in binary FP, except
    where the far-field tail of the wavefront passed through FP32 subnormals (< 1.2e-38)
    and rounded differently; those values stay below 1e-6 of the peak."""
    c = case("ram_fs0", nt=150)
    R1, F1, _ = _run(c)
    R2, F2, _ = _run(c, scale=2.0)
    np.testing.assert_array_equal(R2, 2 * R1)
    big = np.abs(F1) > 1e-6 * np.abs(F1).max()
    np.testing.assert_array_equal(F2[big], 2 * F1[big])
    assert rel_l2(F2, 2 * F1) < 1e-12


def test_oracle_zero_in_zero_out():
    c = case("ram_fs1", nt=80)
    R, F, k = _run(c, scale=0.0)
    assert not R.any() and not F.any() and not k.any()


def _mirror_case(variant):
    c = case("disk_fs0", nt=160, variant=variant)
    nx = c["nx"]
    half = c["ctmp"][:, : nx // 2]
    c["ctmp"] = np.asfortranarray(np.concatenate([half, half[:, ::-1]], axis=1))
    c["nxSou"] = np.array([nx // 2], np.int32)       # not centred on a cell: use a pair
    c["nxSou"] = np.array([nx // 2, nx // 2 + 1], np.int32)
    c["nySou"] = np.array([12, 12], np.int32)
    c["S"] = np.asfortranarray(np.repeat(c["S"][:, :1], 2, axis=1))
    rx = np.arange(2, nx // 2 - 1, 2)
    c["nxRec"] = np.concatenate([rx, nx + 1 - rx[::-1]]).astype(np.int32)
    c["nyRec"] = np.full(c["nxRec"].size, 3, np.int32)
    return c


def test_oracle_mirror_symmetry_disk_variant():
    """A model and geometry mirrored about the centre give mirrored traces, bit for bit,
    in the _disk form (full-width velocity pads)."""
    c = _mirror_case(1)
    R, _, _ = _run(c)
    np.testing.assert_array_equal(R, R[:, ::-1])


def test_oracle_ram_variant_breaks_mirror_symmetry():
    """The _ram padding quirk (top/bottom pads only over the first dn columns) is not
    mirror-symmetric: the same mirrored setup gives asymmetric traces."""
    c = _mirror_case(0)
    R, _, _ = _run(c)
    assert not np.array_equal(R, R[:, ::-1])


def test_aligned_decimated_imaging_equals_masked_reference_imaging():
    """The engine's decimated pairing (backward step k images with the snapshot of forward
    step nt+2-k, every store_snap-th) has no reference counterpart.  It equals, exactly in
    FP64 for store_snap = 4 (a power of two), 4 x the reference-bookkeeping imaging at
    store_snap = 1 with every snapshot the decimated forward would not store set to zero."""
    c = case("ram_fs0", nt=201)
    geo = (c["nxSou"], c["nySou"], c["nxRec"], c["nyRec"])
    nt, ss = c["nt"], 4
    R, F1 = O.forward(*geo, c["S"], c["dx"], c["dt"], c["ctmp"], 1, nt, nt, c["dPML"], 0, precision="f64")
    Sadj = np.asfortranarray(R[::-1, :].astype(np.float32))
    # decimated forward stores F(s') = F1(ss*s'), s' = 1..(nt-1)//ss
    nsnap = (nt - 1) // ss
    Rd, Fd = O.forward(*geo, c["S"], c["dx"], c["dt"], c["ctmp"], ss, nsnap, nt, c["dPML"], 0, precision="f64")
    np.testing.assert_array_equal(Fd, F1[:, :, ss - 1:ss * nsnap:ss])
    k_aligned = O.backward(*geo, Sadj, c["dx"], c["dt"], c["ctmp"], ss, nsnap, nt, c["dPML"], 0, Fields=Fd,
                           variant=2, precision="f64")
    Fm = np.zeros_like(F1)
    Fm[:, :, ss - 1:ss * nsnap:ss] = F1[:, :, ss - 1:ss * nsnap:ss]
    k_masked = O.backward(*geo, Sadj, c["dx"], c["dt"], c["ctmp"], 1, nt, nt, c["dPML"], 0, Fields=Fm,
                          precision="f64")
    assert np.linalg.norm(k_masked) > 0
    np.testing.assert_allclose(k_aligned, ss * k_masked, rtol=1e-13, atol=1e-13 * np.abs(k_masked).max())


def test_fp64_accuracy_fixture_is_the_oracle():
    """The FP64 image / traces fixture the short accuracy figures are stated against is what
    the FP64 oracle computes for its workload (regenerated here with the fixture's own
    generator and compared)."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_accuracy_fixture as MF

    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "accuracy_single401_fp64.npz"))
    d = MF.observed("single401")
    assert float(np.sum(d.astype(np.float64) ** 2)) == float(fx["data_sumsq"])
    k64, R64 = MF.fp64("single401", d)
    np.testing.assert_array_equal(k64.astype(np.float32), fx["image_fp64"])
    np.testing.assert_array_equal(R64[:, fx["trace_channels"]].astype(np.float32), fx["traces_fp64"])
    assert np.abs(fx["image_fp64"]).max() > 0


def test_fp64_full_record_fixture_matches_its_workload():
    """The full-record mcs2001 fixture (minutes to regenerate, so not redone here) belongs
    to shot 0 of the line at nt 6000, holds a non-zero image of the window's shape and the
    traces of every 10th channel; its gather sample is the FP32 oracle's first samples'
    shape (the GPU bench confirms the whole gather by its sum of squares)."""
    from psdm_amd.workloads import line_workload

    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "accuracy_mcs2001_fp64.npz"))
    W = line_workload("mcs2001")
    sh = W["shots"][0]
    assert (int(fx["nt"]), int(fx["x_i"]), int(fx["nx_w"])) == (W["nt"], sh.x_i, sh.nx_w)
    assert fx["image_fp64"].shape == (W["model"].nz, sh.nx_w) and np.abs(fx["image_fp64"]).max() > 0
    np.testing.assert_array_equal(fx["trace_channels"], np.arange(0, sh.nxRec.size, 10))
    assert fx["traces_fp64"].shape == (W["nt"], fx["trace_channels"].size)
    assert bool(fx["direct_wave_removed"]) and float(fx["data_sumsq"]) > 0
