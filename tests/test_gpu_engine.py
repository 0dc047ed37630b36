"""GPU tests of the batched / served engine (psdm_amd.engine -> libpsdm_rtm.so).

* FP32, store_snap = 1: every shot's image is bit-identical to the oracle's
  solver_ac_forward_ram + solver_ac_backward_ram on that shot's window, whatever else
  shares the engine, and the FP64 stack is the sum of those images.
* A shot gives the same image alone and in a mix (bit for bit), also with 16-bit
  snapshots; graph replay and eager launches agree bit for bit.
* HBM-budget admission serialises shots that do not fit together.
* Mixed precision (FP16 / BF16 snapshots every 4th step, Kahan image) is measured
  against the FP64 oracle of the same decimated algorithm; bars in the test.
"""
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


NY, NXP, DPML, DX, DT = 70, 160, 10, 10.0, 0.001


def model(seed=0):
    rng = np.random.default_rng(seed)
    v = 1500 + 1200 * rng.random((NY, NXP))
    k = np.ones(5) / 5
    v = np.apply_along_axis(lambda c: np.convolve(c, k, mode="same"), 1, v)
    v[:8, :] = 1500.0
    return np.asfortranarray(v.astype(np.float32))


def wavelet(nt, f0=25.0):
    t = np.arange(nt) * DT
    a = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


# (x_i, nx_w, nt, source column inside the window)
SHOTS = [(1, 60, 260, 12), (31, 70, 300, 20), (71, 80, 220, 35), (101, 56, 280, 28), (11, 64, 240, 30)]


def shot_inputs(vel, sh, seed):
    x_i, nx_w, nt, sx = sh
    rng = np.random.default_rng(100 + seed)
    nxR = np.arange(3, nx_w - 2, 3, dtype=np.int32)
    nyR = np.full(nxR.size, 2, np.int32)
    data = np.asfortranarray((rng.standard_normal((nt, nxR.size)) * 1e-2).astype(np.float32))
    return dict(x_i=x_i, nx_w=nx_w, nt=nt, nxSou=sx, nySou=3, nxRec=nxR, nyRec=nyR, data=data,
                ctmp=np.asfortranarray(vel[:, x_i - 1:x_i - 1 + nx_w]), S=wavelet(nt))


def oracle_image(si, ss=1, precision="f32"):
    nt = si["nt"]
    geo = ([si["nxSou"]], [si["nySou"]], si["nxRec"], si["nyRec"])
    if ss == 1:
        Rec, F = O.forward(*geo, si["S"].reshape(nt, 1), DX, DT, si["ctmp"], 1, nt, nt, DPML, 0, table_mode=0,
                           precision=precision)
        return O.backward(*geo, np.asfortranarray(si["data"][::-1, :]), DX, DT, si["ctmp"], 1, nt, nt, DPML, 0,
                          Fields=F, table_mode=0, precision=precision)
    nsnap = (nt - 1) // ss
    Rec, F = O.forward(*geo, si["S"].reshape(nt, 1), DX, DT, si["ctmp"], ss, nsnap, nt, DPML, 0, table_mode=0,
                       precision=precision)
    return O.backward(*geo, np.asfortranarray(si["data"][::-1, :]), DX, DT, si["ctmp"], ss, nsnap, nt, DPML, 0,
                      Fields=F, table_mode=0, variant=2, precision=precision)


def run_engine(ctx, vel, shots, idx, nslots, **kw):
    import torch

    from psdm_amd.engine import Engine

    dev = torch.device("cuda:0")
    d_vel = torch.from_numpy(np.ascontiguousarray(vel.T)).to(dev)
    d_sum = torch.zeros((NXP, NY), dtype=torch.float64, device=dev)
    d_fold = torch.zeros((NXP, NY), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    eng = Engine(ctx, d_vel, NY, NXP, nslots=nslots, max_nx=max(s[1] for s in SHOTS),
                 max_nt=max(s[2] for s in SHOTS), dPML=DPML, fs=0, dmodel=DX, dt=DT, d_psdm_sum=d_sum,
                 d_cov_sum=d_fold, **kw)
    outs, keep = {}, []
    for q in idx:
        si = shot_inputs(vel, shots[q], q)
        d_w = torch.from_numpy(si["S"]).to(dev)
        d_d = torch.from_numpy(np.ascontiguousarray(si["data"].T)).to(dev)
        d_img = torch.zeros((si["nx_w"], NY), device=dev)
        torch.cuda.synchronize()
        eng.submit(si["x_i"], si["nx_w"], si["nt"], si["nxSou"], si["nySou"], si["nxRec"], si["nyRec"], d_w, d_d,
                   d_img)
        outs[q] = d_img
        keep += [d_w, d_d]
    eng.drain()
    done = {}
    for q, d in outs.items():
        done[q] = d.cpu().numpy().T.copy(order="F")
    stack = d_sum.cpu().numpy().T, d_fold.cpu().numpy().T
    st = eng.stats()
    eng.close()
    return done, stack, st


def test_engine_fp32_bit_identical_to_oracle_and_stack(ctx):
    vel = model(0)
    say("[gpu] engine fp32: 5 shots, 2 slots, chunk 16")
    imgs, (psum, fold), st = run_engine(ctx, vel, SHOTS, range(5), nslots=2, chunk_steps=16)
    ref_sum = np.zeros((NY, NXP))
    ref_fold = np.zeros((NY, NXP), np.int32)
    for q in range(5):
        si = shot_inputs(vel, SHOTS[q], q)
        ko = oracle_image(si)
        np.testing.assert_array_equal(imgs[q], ko)
        ref_sum[:, si["x_i"] - 1:si["x_i"] - 1 + si["nx_w"]] += ko.astype(np.float64)
        ref_fold[:, si["x_i"] - 1:si["x_i"] - 1 + si["nx_w"]] += 1
    np.testing.assert_array_equal(fold, ref_fold)
    np.testing.assert_allclose(psum, ref_sum, rtol=1e-12, atol=1e-12 * np.abs(ref_sum).max())
    assert st["hbm_used"] == 0 and st["slot_steps"] == sum(2 * (s[2] - 1) for s in SHOTS)


@pytest.mark.parametrize("fmt", ["f32", "f16"])
def test_engine_shot_alone_equals_shot_in_mix(ctx, fmt):
    vel = model(1)
    kw = dict(snap_format=fmt, kahan=fmt != "f32", store_snap=1 if fmt == "f32" else 4, chunk_steps=32)
    say(f"[gpu] engine alone-vs-mix fmt={fmt}")
    alone, _, _ = run_engine(ctx, vel, SHOTS, [2], nslots=1, **kw)
    mixed, _, _ = run_engine(ctx, vel, SHOTS, [0, 1, 2, 3, 4], nslots=3, **kw)
    np.testing.assert_array_equal(alone[2], mixed[2])
    assert np.linalg.norm(alone[2]) > 0


def test_engine_graph_equals_eager_and_grid_shapes(ctx):
    vel = model(2)
    say("[gpu] engine graph vs eager vs persistent grid vs tile_x 32")
    a, _, _ = run_engine(ctx, vel, SHOTS, [0, 1, 3], nslots=2, use_graph=True, chunk_steps=24)
    b, _, _ = run_engine(ctx, vel, SHOTS, [0, 1, 3], nslots=2, use_graph=False, chunk_steps=7)
    c, _, _ = run_engine(ctx, vel, SHOTS, [0, 1, 3], nslots=3, grid=64, tile_x=16, chunk_steps=40)
    d, _, _ = run_engine(ctx, vel, SHOTS, [0, 1, 3], nslots=2, tile_x=32, chunk_steps=64)
    for q in (0, 1, 3):
        np.testing.assert_array_equal(a[q], b[q])
        np.testing.assert_array_equal(a[q], c[q])
        np.testing.assert_array_equal(a[q], d[q])


def test_engine_launch_chains_bit_identical(ctx):
    """Slots split into 1, 2 or 3 concurrent launch chains per captured chunk (5 slots:
    uneven 2+2+1 split with 3) give the same images and stack, bit for bit."""
    vel = model(5)
    say("[gpu] engine launch chains 1 / 2 / 3")
    runs = [run_engine(ctx, vel, SHOTS, range(5), nslots=5, groups=g, chunk_steps=16) for g in (1, 2, 3)]
    for imgs, (psum, fold), _ in runs[1:]:
        for q in range(5):
            np.testing.assert_array_equal(imgs[q], runs[0][0][q])
        np.testing.assert_array_equal(psum, runs[0][1][0])
        np.testing.assert_array_equal(fold, runs[0][1][1])


def test_engine_compact_lookup_paths_bit_identical(ctx):
    """70 resident shots of uneven windows: one group of 70 slots (the slot of a work item
    found by a walk over the slot descriptors) and two groups of 35 (a wave-wide compare of
    the group's offsets) give every shot the image it gets in a 5-slot engine."""
    vel = model(6)
    idx = [q % 5 for q in range(70)]
    say("[gpu] engine 70 slots: 1 group of 70 / 2 groups of 35 / reference 5 slots")
    ref, (_, fold_ref), _ = run_engine(ctx, vel, SHOTS, range(5), nslots=5, chunk_steps=32)
    for groups in (1, 2):
        imgs, (_, fold), _ = run_engine(ctx, vel, SHOTS, idx, nslots=70, groups=groups, chunk_steps=32)
        for q in range(5):
            np.testing.assert_array_equal(imgs[q], ref[q])
        np.testing.assert_array_equal(fold, 14 * fold_ref)


def test_engine_hbm_budget_admission(ctx):
    """A budget that holds one FP32 snapshot store at a time runs the shots one after
    the other in a 3-slot engine, with the same results."""
    vel = model(3)
    lds = -(-(max(NY, 248) + 8) // 64) * 64  # snapshot column length (interior_ld)
    one = (SHOTS[1][2] - 1) * lds * SHOTS[1][1] * 4 + 4096
    say("[gpu] engine hbm budget admission")
    a, _, sa = run_engine(ctx, vel, SHOTS, [1, 1, 1], nslots=3, hbm_budget=one, chunk_steps=16)
    b, _, sb = run_engine(ctx, vel, SHOTS, [1], nslots=1, chunk_steps=16)
    np.testing.assert_array_equal(a[1], b[1])
    # serialised: three shots take three times the steps of one
    assert sa["steps_issued"] >= 3 * 2 * (SHOTS[1][2] - 1)


def test_one_slot_arena_holds_a_full_size_16_bit_shot(ctx):
    """A one-slot engine sized for exactly one shot of the largest window and record, with
    FP16 snapshots every 4th step: the footprint (snapshots plus one FP32 scale per
    248-row segment of each snapshot column, 65 x 57 x 4 bytes here) is not a multiple of
    256 bytes, and the arena hands out 256-byte granules.  The default budget counts
    whole granules, so the shot is admitted and gives the image it gives in a two-slot
    engine (regression: it was never admitted and the drain did not return)."""
    import torch

    from psdm_amd.engine import Engine

    vel = model(4)
    nx_w, nt = 57, 261
    assert ((nt - 1) // 4 * nx_w * 4) % 256 != 0
    si = shot_inputs(vel, (21, nx_w, nt, 20), 9)
    dev = torch.device("cuda:0")
    d_vel = torch.from_numpy(np.ascontiguousarray(vel.T)).to(dev)
    d_w = torch.from_numpy(si["S"]).to(dev)
    d_d = torch.from_numpy(np.ascontiguousarray(si["data"].T)).to(dev)
    imgs = []
    for nslots in (1, 2):
        say(f"[gpu] one full-size fp16 shot, {nslots} slot(s)")
        eng = Engine(ctx, d_vel, NY, NXP, nslots=nslots, max_nx=nx_w, max_nt=nt, dPML=DPML, fs=0, dmodel=DX, dt=DT,
                     snap_format="f16", store_snap=4, kahan=True, chunk_steps=16)
        d_img = torch.zeros((nx_w, NY), device=dev)
        torch.cuda.synchronize()
        eng.submit(si["x_i"], nx_w, nt, si["nxSou"], si["nySou"], si["nxRec"], si["nyRec"], d_w, d_d, d_img)
        eng.drain()
        st = eng.stats()
        assert st["hbm_budget"] % 256 == 0
        imgs.append(d_img.cpu().numpy().T.copy(order="F"))
        eng.close()
    assert np.linalg.norm(imgs[0]) > 0
    np.testing.assert_array_equal(imgs[0], imgs[1])


def test_engine_rejects_oversized_and_out_of_model(ctx):
    import torch

    from psdm_amd._native import RtmError
    from psdm_amd.engine import Engine

    vel = model(0)
    dev = torch.device("cuda:0")
    d_vel = torch.from_numpy(np.ascontiguousarray(vel.T)).to(dev)
    eng = Engine(ctx, d_vel, NY, NXP, nslots=1, max_nx=60, max_nt=200, dPML=DPML, fs=0, dmodel=DX, dt=DT)
    w = torch.zeros(300, device=dev)
    d = torch.zeros((3, 300), device=dev)
    with pytest.raises(RtmError, match="wider"):
        eng.submit(1, 70, 200, 5, 3, [3, 6, 9], [2, 2, 2], w, d)
    with pytest.raises(RtmError, match="record length"):
        eng.submit(1, 50, 300, 5, 3, [3, 6, 9], [2, 2, 2], w, d)
    with pytest.raises(RtmError, match="outside the velocity model"):
        eng.submit(140, 40, 200, 5, 3, [3, 6, 9], [2, 2, 2], w, d)
    eng.close()


def test_engine_out_of_memory_leaves_no_stale_error(ctx):
    """A snapshot arena larger than the device fails engine creation with an error, and
    the caller's next launch (torch checks the thread's last HIP error after it) works."""
    import torch

    from psdm_amd._native import RtmError
    from psdm_amd.engine import Engine

    dev = torch.device("cuda:0")
    d_vel = torch.from_numpy(np.ascontiguousarray(model(0).T)).to(dev)
    with pytest.raises(RtmError, match="out of memory"):
        Engine(ctx, d_vel, NY, NXP, nslots=1, max_nx=60, max_nt=200, dPML=DPML, fs=0, dmodel=DX, dt=DT,
               hbm_budget=1 << 44)
    x = torch.zeros((NXP, NY), dtype=torch.float64, device=dev) + 1.0
    torch.cuda.synchronize()
    assert float(x.sum()) == NXP * NY


@pytest.mark.parametrize("fmt,bar", [("f32", 1e-5), ("f16", 2e-4), ("bf16", 2e-3)])
def test_engine_mixed_precision_vs_fp64_oracle(ctx, fmt, bar):
    """Image of a decimated (store_snap 4) shot vs the FP64 oracle of the same algorithm:
    FP32 snapshots stay at FP32 rounding level; FP16 / BF16 within the stated bars."""
    vel = model(4)
    say(f"[gpu] engine mixed precision fmt={fmt}")
    imgs, _, _ = run_engine(ctx, vel, SHOTS, [1], nslots=1, snap_format=fmt, kahan=True, store_snap=4,
                            chunk_steps=32)
    si = shot_inputs(vel, SHOTS[1], 1)
    k64 = oracle_image(si, ss=4, precision="f64")
    err = rel_l2(imgs[1], k64)
    say(f"  {fmt}: image rel-L2 vs FP64 oracle = {err:.3e}")
    assert err <= bar
