"""The depth-to-time oracle (oracle/poststack.py) pinned against the reference's own
linear_interpolation_1D (compiled into oracle/_ref) and against closed forms."""
import numpy as np
import pytest

from oracle import poststack as P
from oracle import pyoracle

need_ref = pytest.mark.skipif(not pyoracle.ref_available(), reason="oracle/_ref not built")


@need_ref
@pytest.mark.parametrize("seed", range(12))
def test_linear_interpolation_bit_exact_vs_reference(seed):
    """Monotone abscissae with flat runs (zero velocity), zero ordinates (skipped
    intervals), non-monotone abscissae, abscissae on breakpoints and beyond both ends."""
    rng = np.random.default_rng(seed)
    n1 = int(rng.integers(2, 80))
    inc = rng.uniform(0, 0.01, n1).astype(np.float32)
    inc[rng.random(n1) < 0.2] = 0
    x1 = np.add.accumulate(inc).astype(np.float32)
    if seed % 4 == 3:
        x1 = rng.uniform(0, 0.3, n1).astype(np.float32)
    y1 = rng.standard_normal(n1).astype(np.float32)
    y1[rng.random(n1) < 0.25] = 0
    x2 = np.concatenate([np.arange(int(rng.integers(1, 120))) * np.float32(0.004),
                         x1[rng.integers(0, n1, 6)], [-1.0, 9.0]]).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        np.testing.assert_array_equal(P.linear_interpolation_1d(x1, y1, x2), P.ref_linear_interpolation_1d(x1, y1, x2))


def test_linear_interpolation_closed_form():
    """Samples of a line with no zero ordinates are reproduced on the line."""
    x1 = np.linspace(0, 1, 11).astype(np.float32)
    y1 = (2.0 + 3.0 * x1).astype(np.float32)
    x2 = np.linspace(-0.5, 1.5, 41).astype(np.float32)
    y = P.linear_interpolation_1d(x1, y1, x2)
    want = np.clip(2.0 + 3.0 * x2, 2.0, 5.0)
    np.testing.assert_allclose(y, want, rtol=0, atol=2e-6)


def test_twt_times_closed_form_and_order():
    """Constant velocity: t(k) = (k-1) * 2 dd / v up to FP32 accumulation, accumulated in
    depth order; zero velocity adds nothing; nt = 1 + ceiling(max t / dt)."""
    nz, dd, v = 400, 10.0, 1500.0
    vel = np.full((nz, 3), v, np.float32)
    vel[100:, 1] = 3000.0
    vel[50:60, 2] = 0.0
    t = P.twt_times(vel, dd)
    np.testing.assert_allclose(t[:, 0], np.arange(nz) * 2 * dd / v, rtol=2e-5)
    inc = np.float32(np.float32(2 * dd) / np.float32(v))
    seq = np.float32(0)
    for k in range(1, nz):
        seq = np.float32(seq + inc)
    assert t[-1, 0] == seq
    np.testing.assert_allclose(t[-1, 1], 99 * 2 * dd / v + 300 * 2 * dd / 3000.0, rtol=2e-5)
    assert np.all(t[50:60, 2] == t[49, 2])
    assert P.twt_length(t, 0.004) == 1 + int(np.ceil(np.float32(t.max()) / np.float32(0.004)))


@need_ref
@pytest.mark.parametrize("seed", range(6))
def test_normalize_f32_bit_exact_vs_reference(seed):
    """normalize_f32 against the reference's normalize_psdm: folds 0, 1 and > 1, negative
    and zero sums, a fold that is not an integer (the reference divides by any positive
    FP32 Cov_sum)."""
    rng = np.random.default_rng(seed)
    ny, nx = int(rng.integers(1, 90)), int(rng.integers(1, 70))
    s = (rng.standard_normal((ny, nx)) * 10.0 ** rng.integers(-6, 6)).astype(np.float32)
    c = rng.integers(-2, 40, (ny, nx)).astype(np.float32)
    c[rng.random((ny, nx)) < 0.1] = np.float32(0.5)
    s[rng.random((ny, nx)) < 0.1] = 0
    np.testing.assert_array_equal(P.normalize_f32(s, c), P.ref_normalize_psdm(s, c))


@pytest.mark.parametrize("seed", range(6))
def test_normalize_fp64_stack_equals_fp32_division(seed):
    """The FP64-stack normalize on an FP32-representable sum and an integer fold is the
    reference's FP32 division bit for bit (division is correctly rounded and FP64 carries
    more than 2 * 24 + 2 significand bits); zero where the fold is zero."""
    rng = np.random.default_rng(100 + seed)
    shape = (257, 131)
    s32 = (rng.standard_normal(shape) * 10.0 ** rng.integers(-30, 30, shape)).astype(np.float32)
    fold = rng.integers(0, 5000, shape).astype(np.int32)
    fold[rng.random(shape) < 0.2] = 0
    fold[rng.random(shape) < 0.2] = 1
    got = P.normalize(s32.astype(np.float64), fold)
    np.testing.assert_array_equal(got, P.normalize_f32(s32, fold.astype(np.float32)))
    assert np.all(got[fold == 0] == 0)
    np.testing.assert_array_equal(got[fold == 1], s32[fold == 1])


def test_normalize_rounds_the_fp64_quotient_once():
    """A sum that FP32 cannot hold: the result is the FP64 quotient rounded once (the FP32
    stack of the reference would have rounded the sum first)."""
    s = np.array([[1.0 + 2.0 ** -40, 3.0 * (1 + 2.0 ** -30)]])
    fold = np.array([[3, 3]], np.int32)
    got = P.normalize(s, fold)
    np.testing.assert_array_equal(got, (s / 3.0).astype(np.float32))
