"""GPU parity of the gather conditioning (psdm_rtm_condition) with the reference's
routines (golden fixture written from oracle/_ref, and oracle/_ref itself where it
travels with the tree) and with the oracle restatement (oracle/conditioning.py).

Bars: resampling, cuts and tapers bit for bit; the filter within 1e-5 rel-L2 of the
reference's FFTPACK evaluation (the device applies the same operator as a convolution
with FP64-designed taps and FP32 accumulation)."""
import os
import sys

import numpy as np
import pytest

from oracle import conditioning as O
from oracle import pyoracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FILTER_BAR = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(ctx, raw, **kw):
    """Device conditioning of raw(nt, n) (Fortran order) -> out(nt_taper, n)."""
    import torch

    from psdm_amd.conditioning import condition

    d = torch.from_numpy(np.ascontiguousarray(np.asarray(raw, np.float32).T)).to("cuda:0")
    out = condition(ctx, d, **kw)
    torch.cuda.synchronize()
    return np.asfortranarray(out.cpu().numpy().T)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_golden_fixture_cases(ctx):
    z = np.load(os.path.join(HERE, "golden", "conditioning_ref.npz"))
    for name in [k[:-4] for k in z.files if k.endswith("_raw")]:
        p = dict(zip(z["param_names"], z[name + "_params"]))
        say(f"[gpu] conditioning golden case {name}")
        out = run(ctx, z[name + "_raw"], dt_data=float(p["dt_data"]), typef=int(p["typef"]),
                  f_init=float(p["f_init"]), f_final=float(p["f_final"]), nt_prop2=int(p["nt_prop2"]),
                  dt_prop=float(p["dt_prop"]), nt_taper=int(p["nt_taper"]), N_taper=int(p["N_taper"]),
                  N_zeros=int(p["N_zeros"]), taper_option=int(p["taper_option"]),
                  nt_sf=z[name + "_nt_sf"] if name + "_nt_sf" in z.files else None,
                  nt_cut=z[name + "_nt_cut"] if name + "_nt_cut" in z.files else None)
        ref = z[name + "_out"]
        if float(p["f_init"]) > 0:
            assert rel(out, ref) < FILTER_BAR, name
            assert not np.any(out[:, -1]), name  # the all-zero trace stays zero
        else:
            np.testing.assert_array_equal(out, ref, err_msg=name)


AKIMA = [  # (L, n, dt_data, nt, dt): up/down-sampling, outputs past X(L), tiny L
    (500, 7, 0.002, 1200, 0.0008), (900, 3, 0.0008, 330, 0.002), (2, 2, 0.01, 5, 0.004), (3, 2, 0.01, 9, 0.004),
    (4, 2, 0.01, 12, 0.004), (60, 70, 0.004, 200, 0.0016)]


@pytest.mark.parametrize("case", AKIMA)
def test_resample_bit_exact_vs_oracle(ctx, case):
    L, n, dtd, nt, dt = case
    say(f"[gpu] conditioning akima L={L} n={n} nt={nt}")
    rng = np.random.default_rng(L)
    raw = np.asfortranarray(rng.standard_normal((L, n)).astype(np.float32))
    raw[:, 0] = np.sin(0.3 * np.arange(L))
    if L > 4:
        raw[:L // 2, 1] = 0.0
        raw[L // 2:, 1] = 1.0
    out = run(ctx, raw, dt_data=dtd, nt_prop2=nt, dt_prop=dt)
    np.testing.assert_array_equal(out, O.akima(raw, dtd, nt, dt))
    if pyoracle.ref_available():
        np.testing.assert_array_equal(out, O.ref_akima(raw, dtd, nt, dt))


def test_cuts_and_tapers_bit_exact_vs_oracle(ctx):
    rng = np.random.default_rng(5)
    L, n = 400, 9
    raw = np.asfortranarray(rng.standard_normal((L, n)).astype(np.float32))
    kw = dict(dt_data=0.002, nt_prop2=950, dt_prop=0.0008, nt_taper=1000)
    nt_sf = rng.integers(0, 400, n).astype(np.int32)
    nt_cut = np.array([1, 2, 10, 100, 500, 949, 950, 999, 1000], np.int32)
    for mode in ("taper", "cut", "taper_short"):
        say(f"[gpu] conditioning cuts/tapers mode={mode}")
        if mode == "taper":
            a = run(ctx, raw, nt_sf=nt_sf, taper_option=1, N_taper=120, N_zeros=50, **kw)
            b = O.condition(raw, typef=3, f_init=0, f_final=0, nt_sf=nt_sf, taper_option=1, N_taper=120,
                            N_zeros=50, **kw)
        elif mode == "cut":
            a = run(ctx, raw, nt_cut=nt_cut, N_taper=37, N_zeros=0, **kw)
            b = O.condition(raw, typef=3, f_init=0, f_final=0, nt_cut=nt_cut, N_taper=37, N_zeros=0, **kw)
        else:  # taper longer than the kept samples
            kw2 = dict(kw, nt_prop2=20, nt_taper=40)
            a = run(ctx, raw, taper_option=1, N_taper=64, N_zeros=20, **kw2)
            b = O.condition(raw, typef=3, f_init=0, f_final=0, taper_option=1, N_taper=64, N_zeros=20, **kw2)
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("typef,f1,f2", [(3, 5.0, 40.0), (1, 3.0, 25.0), (2, 5.0, 60.0)])
def test_filter_vs_reference_full_size(ctx, typef, f1, f2):
    """A 2401-sample, 96-trace gather at 2 ms (the raw record of a 4.8 s MCS shot)."""
    nt, n, dt = 2401, 96, 0.002
    rng = np.random.default_rng(typef)
    t = np.arange(nt) * dt
    raw = np.zeros((nt, n), np.float32, order="F")
    for i in range(n):
        for t0 in rng.uniform(0.1, 4.5, 4):
            a = (np.pi * 25.0 * (t - t0)) ** 2
            raw[:, i] += rng.uniform(-1, 1) * (1 - 2 * a) * np.exp(-a)
        raw[:, i] += 0.02 * rng.standard_normal(nt)
    raw[:, 5] = 0.0
    say(f"[gpu] conditioning filter typef={typef} {nt}x{n}")
    out = run(ctx, raw, dt_data=dt, typef=typef, f_init=f1, f_final=f2, nt_prop2=nt, dt_prop=dt)
    orc = O.freq_filter(raw, dt, typef, f1, f2)
    assert rel(out, orc) < FILTER_BAR
    assert not np.any(out[:, 5])
    if pyoracle.ref_available():
        assert rel(out, O.ref_freq_filter(raw, dt, typef, f1, f2)) < FILTER_BAR


def test_filter_cache_and_full_pipeline(ctx):
    """Alternating filters (the taps are cached per filter) and the whole driver
    sequence against the oracle: filter, resample 2 ms -> 0.8 ms, cut_sf, taper."""
    rng = np.random.default_rng(9)
    nt, n = 800, 33
    raw = np.asfortranarray(rng.standard_normal((nt, n)).astype(np.float32))
    kw = dict(dt_data=0.002, nt_prop2=1990, dt_prop=0.0008, nt_taper=2050, taper_option=1, N_taper=50, N_zeros=60)
    nt_sf = rng.integers(0, 300, n).astype(np.int32)
    outs = []
    for rep, (typef, f1, f2) in enumerate([(3, 5.0, 40.0), (1, 5.0, 30.0), (3, 5.0, 40.0)]):
        say(f"[gpu] conditioning pipeline rep={rep} typef={typef}")
        a = run(ctx, raw, typef=typef, f_init=f1, f_final=f2, nt_sf=nt_sf, **kw)
        b = O.condition(raw, typef=typef, f_init=f1, f_final=f2, nt_sf=nt_sf, **kw)
        assert rel(a, b) < FILTER_BAR
        outs.append(a)
    np.testing.assert_array_equal(outs[0], outs[2])


def test_rejects_bad_arguments(ctx):
    import torch

    from psdm_amd._native import RtmError
    from psdm_amd.conditioning import condition

    d = torch.zeros((4, 100), device="cuda:0")
    with pytest.raises(RtmError, match="nt_prop2 > nt_data"):
        condition(ctx, d, dt_data=0.002, nt_prop2=150, dt_prop=0.002)
    with pytest.raises(RtmError, match="nt_cut"):
        condition(ctx, d, dt_data=0.002, nt_prop2=80, dt_prop=0.002, nt_taper=90, N_taper=5,
                  nt_cut=[1, 2, 91, 4])
    with pytest.raises(RtmError, match="typef"):
        condition(ctx, d, dt_data=0.002, nt_prop2=80, dt_prop=0.002, typef=4, f_init=1.0, f_final=2.0)
    with pytest.raises(RtmError, match="N_taper"):
        condition(ctx, d, dt_data=0.002, nt_prop2=80, dt_prop=0.002, taper_option=1, N_taper=0)
