"""GPU tests of the batched engine at production window shapes (psdm_amd.engine).

The shapes are those of the uneven 256-shot line (601 rows @ 10 m, windows of about 600,
1300 and 2000 columns, dPML 40) and of the long-offset WAS gathers (801 x 5102 window,
1000 receivers by reciprocity); record lengths are cut to a few hundred steps so that the
oracle covers them in seconds (the cost and the arithmetic of a step do not depend on the
record length).

* FP32, snapshots every step: every image is bit-identical to the oracle's
  solver_ac_forward_ram + solver_ac_backward_ram on that window, with the three widths
  resident together (compact work list, one and two launch chains), and through one group
  of 66 slots (a work item's slot found by a walk over the descriptors).
* FP16 / BF16 snapshots every 4th step with Kahan imaging, against the FP64 oracle of the
  same decimated algorithm, with the bars stated in the test.
* A WAS window through the engine: FP32 image equal to the oracle's; FP32 and FP16
  images the same alone and in a mix with the MCS windows.
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import pyoracle as O
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

DPML, DX, DT = 40, 10.0, 0.0008


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_LINE = {}


def line():
    """The uneven line's model and its narrowest, median and widest shots."""
    if not _LINE:
        from psdm_amd.workloads import uneven_line

        W = uneven_line(256, seed=0)
        order = sorted(range(len(W["shots"])), key=lambda k: W["shots"][k].nx_w)
        _LINE.update(W=W, pick=[order[0], order[len(order) // 2], order[-1]])
    return _LINE["W"], _LINE["pick"]


def shot_job(m, sh, nt, seed, ricker_delay=None):
    """Inputs of one shot (window-relative geometry), with a seeded noise gather."""
    from psdm_amd.solvers import ricker_mcs

    rng = np.random.default_rng(seed)
    nxR = (sh.nxRec - sh.x_i + 1).astype(np.int32)
    data = np.asfortranarray((rng.standard_normal((nt, nxR.size)) * 1e-2).astype(np.float32))
    f0 = 15.0 if ricker_delay is None else 8.0
    S = ricker_mcs(nt, DT, f0, 1.2 / f0 if ricker_delay is None else ricker_delay)
    return dict(x_i=sh.x_i, nx_w=sh.nx_w, nt=nt, nxSou=sh.nxSou - sh.x_i + 1, nySou=sh.nySou, nxRec=nxR,
                nyRec=sh.nyRec.astype(np.int32), data=data, S=S,
                ctmp=np.asfortranarray(m.migration[:, sh.x_i - 1:sh.x_f]))


def oracle_image(j, ss=1, precision="f32"):
    nt = j["nt"]
    geo = ([j["nxSou"]], [j["nySou"]], j["nxRec"], j["nyRec"])
    adj = np.asfortranarray(j["data"][::-1, :])
    if ss == 1:
        _, F = O.forward(*geo, j["S"].reshape(nt, 1), DX, DT, j["ctmp"], 1, nt, nt, DPML, 0, table_mode=0,
                         precision=precision)
        return O.backward(*geo, adj, DX, DT, j["ctmp"], 1, nt, nt, DPML, 0, Fields=F, table_mode=0,
                          precision=precision)
    nsnap = (nt - 1) // ss
    _, F = O.forward(*geo, j["S"].reshape(nt, 1), DX, DT, j["ctmp"], ss, nsnap, nt, DPML, 0, table_mode=0,
                     precision=precision)
    return O.backward(*geo, adj, DX, DT, j["ctmp"], ss, nsnap, nt, DPML, 0, Fields=F, table_mode=0, variant=2,
                      precision=precision)


def oracle_images(jobs, **kw):
    """Oracle images of several shots on parallel threads (the C oracle runs without the GIL)."""
    with ThreadPoolExecutor(len(jobs)) as ex:
        return list(ex.map(lambda j: oracle_image(j, **kw), jobs))


def run_engine(ctx, m, jobs, nslots, **kw):
    """Images of `jobs` migrated together by one engine on model m, and the engine stats."""
    import torch

    from psdm_amd.engine import Engine

    dev = torch.device("cuda:0")
    d_vel = torch.from_numpy(np.ascontiguousarray(m.migration.T)).to(dev)
    d_sum = torch.zeros((m.nx, m.nz), dtype=torch.float64, device=dev)
    d_fold = torch.zeros((m.nx, m.nz), dtype=torch.int32, device=dev)
    eng = Engine(ctx, d_vel, m.nz, m.nx, nslots=nslots, max_nx=max(j["nx_w"] for j in jobs),
                 max_nt=max(j["nt"] for j in jobs), dPML=DPML, fs=0, dmodel=DX, dt=DT, d_psdm_sum=d_sum,
                 d_cov_sum=d_fold, **kw)
    outs, keep = [], []
    for j in jobs:
        d_w = torch.from_numpy(j["S"]).to(dev)
        d_d = torch.from_numpy(np.ascontiguousarray(j["data"].T)).to(dev)
        d_img = torch.zeros((j["nx_w"], m.nz), device=dev)
        eng.submit(j["x_i"], j["nx_w"], j["nt"], j["nxSou"], j["nySou"], j["nxRec"], j["nyRec"], d_w, d_d, d_img)
        outs.append(d_img)
        keep += [d_w, d_d]
    eng.drain()
    imgs = [d.cpu().numpy().T.copy(order="F") for d in outs]
    fold = d_fold.cpu().numpy().T
    st = eng.stats()
    eng.close()
    return imgs, fold, st


NTS = (100, 140, 120)  # uneven record lengths of the three windows


def test_uneven_windows_fp32_bit_identical_to_oracle(ctx):
    W, pick = line()
    m = W["model"]
    jobs = [shot_job(m, W["shots"][k], nt, k) for k, nt in zip(pick, NTS)]
    say(f"[gpu] production engine fp32: windows {[j['nx_w'] for j in jobs]} x 601, nt {NTS}")
    ref = oracle_images(jobs)
    for nslots, groups in ((3, 1), (3, 2), (2, 2)):
        say(f"  nslots={nslots} groups={groups}")
        imgs, fold, st = run_engine(ctx, m, jobs, nslots, groups=groups, chunk_steps=32)
        for j, a, b in zip(jobs, imgs, ref):
            assert np.linalg.norm(b) > 0
            np.testing.assert_array_equal(a, b, err_msg=f"window {j['nx_w']}")
        for j in jobs:
            assert np.all(fold[:, j["x_i"] - 1:j["x_i"] - 1 + j["nx_w"]] >= 1)
        assert st["slot_steps"] == sum(2 * (nt - 1) for nt in NTS)


def test_group_of_66_slots_walk_lookup(ctx):
    """66 resident copies of the narrowest window in one launch group (more slots than a
    wavefront has lanes: the slot of a work item is found by a walk), plus the widest
    window in the same group (uneven items per slot): every image equals the oracle's."""
    W, pick = line()
    m = W["model"]
    a = shot_job(m, W["shots"][pick[0]], 60, 1)
    b = shot_job(m, W["shots"][pick[2]], 72, 2)
    say(f"[gpu] production engine 67 slots, 1 group: 66 x {a['nx_w']} + 1 x {b['nx_w']} columns")
    ra, rb = oracle_images([a, b])
    imgs, fold, _ = run_engine(ctx, m, [a] * 66 + [b], 67, groups=1, chunk_steps=16)
    for q in range(66):
        np.testing.assert_array_equal(imgs[q], ra, err_msg=f"copy {q}")
    np.testing.assert_array_equal(imgs[66], rb)
    assert fold[:, a["x_i"] - 1].max() >= 66


@pytest.mark.parametrize("fmt,bar", [("f32", 1e-5), ("f16", 3e-3), ("bf16", 3e-2)])
def test_uneven_windows_mixed_precision_vs_fp64(ctx, fmt, bar):
    """Snapshots every 4th step (backward step k paired with forward step nt+2-k), Kahan
    image accumulation: image rel-L2 against the FP64 oracle of the same algorithm within
    `bar` for each window.  FP32 snapshots: FP32 rounding level.  FP16: a snapshot sample
    carries up to 2^-11 relative rounding (RMS about 2.8e-4), and a short record's image is
    a sum with much cancellation, which raises its relative error several times (measured
    8.0e-4 on the 611-column window, more on the widest window, whose short-record image is
    about 100x smaller than the median window's); BF16 keeps 3 fewer bits (measured 3.9e-3,
    2.3e-3 and 2.1e-2 on the three windows), so its bar is 10x FP16's."""
    W, pick = line()
    m = W["model"]
    jobs = [shot_job(m, W["shots"][k], nt + 1, k) for k, nt in zip(pick, NTS)]  # nt - 1 divisible by 4
    say(f"[gpu] production engine {fmt} every 4th step: windows {[j['nx_w'] for j in jobs]}")
    ref = oracle_images(jobs, ss=4, precision="f64")
    imgs, _, _ = run_engine(ctx, m, jobs, 3, snap_format=fmt, store_snap=4, kahan=True, chunk_steps=32)
    errs = []
    for j, a, b in zip(jobs, imgs, ref):
        assert np.linalg.norm(b) > 0
        errs.append(rel_l2(a, b))
        say(f"  {fmt} window {j['nx_w']}: image rel-L2 vs FP64 = {errs[-1]:.3e} (bar {bar:g})")
    assert max(errs) <= bar


def test_was_window_through_engine_alone_and_in_mix(ctx):
    """An 801 x 5102 WAS window (OBS as source, 1000 shots as receivers, Ricker_was delay
    1.5/f0): the FP32 image equals the oracle's; in FP32 and in FP16 every 4th step with
    Kahan imaging, the image is the same alone and resident with two MCS windows.  The OBS
    is moved up to 30 m depth: over the 60 steps the oracle covers, the source and receiver
    wavefields then meet and the image is not zero (from the seafloor they would need
    about 1250 steps)."""
    from psdm_amd.workloads import was_gathers

    G = was_gathers(1, seed=2)
    mw, sh = G["model"], G["shots"][0]
    nt = 61
    was = shot_job(mw, sh, nt, 7, ricker_delay=1.5 / 8.0)
    was["nySou"] = 1 + int(30.0 / DX)
    assert was["ctmp"].shape == (801, 5102) and was["nxRec"].size == 1000
    say(f"[gpu] WAS window {was['ctmp'].shape} nt {nt} through the engine")
    ref = oracle_image(was)
    # the MCS windows cut from the WAS model so both kinds share one engine (same depth)
    from psdm_amd.workloads import _mcs_shot

    mcs = [shot_job(mw, _mcs_shot(q, x, DX, mw.nx, 480, 12.5, 150.0, 6.0, 8.0, 500.0), n, 20 + q)
           for q, (x, n) in enumerate(((9000.0, 49), (30000.0, 57)))]
    for kw in (dict(), dict(snap_format="f16", store_snap=4, kahan=True)):
        say(f"  {kw or 'fp32'}: alone (1 slot)")
        alone, _, _ = run_engine(ctx, mw, [was], 1, chunk_steps=16, **kw)
        say(f"  {kw or 'fp32'}: in a mix (3 slots)")
        mixed, _, _ = run_engine(ctx, mw, [mcs[0], was, mcs[1]], 3, chunk_steps=16, **kw)
        say(f"  {kw or 'fp32'}: |image| {np.linalg.norm(alone[0]):.3e}")
        np.testing.assert_array_equal(alone[0], mixed[1])
        if not kw:
            np.testing.assert_array_equal(alone[0], ref)
    assert np.linalg.norm(ref) > 0
