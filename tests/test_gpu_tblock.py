"""GPU tests of two time steps per launch (temporal blocking) in the batched engine.

With steps_per_launch = 2 every launch advances each resident shot two time steps, the
first one computed redundantly over a halo (CPML edge tiles in LDS, interior stream items
with a two-level register pipeline).  The per-step arithmetic is unchanged, so every image
must be bit-identical to the one-step engine's and to the oracle's:

* the single-source parity cases (free surface on and off, the _ram and _disk stepping,
  receivers duplicated and outside the interior, several tiles per axis), record lengths
  with an odd and an even number of steps (a phase's last launch then advances one step);
* decimated snapshots (every 2nd / 4th step) with plain and Kahan imaging;
* the uneven five-shot mix in one and two launch chains and with admission mid-run;
* the 601-row production windows (about 600, 1300 and 2000 columns).
"""
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from tests.cases import case

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def engine_image(ctx, c, data, steps_per_launch, store_snap=1, kahan=False, **kw):
    """Image of case c (window = the whole model, one source) migrated by the engine."""
    import torch

    from psdm_amd.engine import Engine

    dev = torch.device("cuda:0")
    ny, nx, nt = c["ny"], c["nx"], c["nt"]
    d_vel = torch.from_numpy(np.ascontiguousarray(c["ctmp"].T)).to(dev)
    eng = Engine(ctx, d_vel, ny, nx, nslots=1, max_nx=nx, max_nt=nt, dPML=c["dPML"], fs=c["fs"],
                 variant=c["variant"], dmodel=c["dx"], dt=c["dt"], store_snap=store_snap, kahan=kahan,
                 steps_per_launch=steps_per_launch, chunk_steps=kw.pop("chunk_steps", 16), **kw)
    d_w = torch.from_numpy(np.ascontiguousarray(c["S"][:, 0])).to(dev)
    d_d = torch.from_numpy(np.ascontiguousarray(data.T)).to(dev)
    d_img = torch.zeros((nx, ny), device=dev)
    eng.submit(1, nx, nt, int(c["nxSou"][0]), int(c["nySou"][0]), c["nxRec"], c["nyRec"], d_w, d_d, d_img)
    eng.drain()
    img = d_img.cpu().numpy().T.copy(order="F")
    eng.close()
    return img


def oracle_case(c, store_snap=1):
    """(traces of the forward modelling, image of the engine's pairing) from the oracle."""
    geo = (c["nxSou"][:1], c["nySou"][:1], c["nxRec"], c["nyRec"])
    nt = c["nt"]
    S = np.asfortranarray(c["S"][:, :1])
    Rec, _ = O.forward(*geo, S, c["dx"], c["dt"], c["ctmp"], 1, 0, nt, c["dPML"], c["fs"], variant=c["variant"],
                       table_mode=0)
    # one snapshot per step: the drivers' nt_snap = nt bookkeeping pairs backward step k with
    # forward step nt + 2 - k, as the engine does; decimated: the engine's aligned pairing
    nsnap = nt if store_snap == 1 else (nt - 1) // store_snap
    _, F = O.forward(*geo, S, c["dx"], c["dt"], c["ctmp"], store_snap, nsnap, nt, c["dPML"], c["fs"],
                     variant=c["variant"], table_mode=0)
    k = O.backward(*geo, np.asfortranarray(Rec[::-1, :]), c["dx"], c["dt"], c["ctmp"], store_snap, nsnap, nt,
                   c["dPML"], c["fs"], Fields=F, variant=c["variant"] | (2 if store_snap > 1 else 0), table_mode=0)
    return Rec, k


# single-source parity cases (the engine takes one source per shot); nt - 1 odd and even
TB_CASES = {
    "ram_fs0": dict(),
    "ram_fs1": dict(fs=1, seed=1),
    "disk_fs1": dict(fs=1, variant=1, seed=2),
    "disk_fs0_even": dict(fs=0, variant=1, seed=3, src=(10, 20), nt=241),
    "ram_big": dict(ny=140, nx=96, nt=220, seed=6, src=(48, 4), rec_y=2),
    "ram_pml_dup_rec": dict(seed=5, rec_x=[5, 5, 9, 9, 9, 14, -3, 60, 25], rec_y=[3, 3, 1, 1, 1, 40, 2, 5, -4]),
    "ram_fs1_even_wide": dict(fs=1, seed=9, ny=64, nx=120, nt=181, src=(60, 4), rec_y=2),
}


@pytest.mark.parametrize("name", list(TB_CASES))
def test_two_steps_equal_one_step_and_oracle(ctx, name):
    base = dict(TB_CASES[name])
    c = case("ram_fs0", **base)
    say(f"[gpu] two steps per launch: case {name} {c['ny']}x{c['nx']} nt={c['nt']} fs={c['fs']} "
        f"variant={c['variant']}")
    Rec, ko = oracle_case(c)
    k1 = engine_image(ctx, c, Rec, 1)
    k2 = engine_image(ctx, c, Rec, 2)
    assert np.linalg.norm(ko) > 0
    np.testing.assert_array_equal(k1, ko)
    np.testing.assert_array_equal(k2, ko)


@pytest.mark.parametrize("ss,kahan", [(2, False), (4, True), (1, True)])
def test_two_steps_decimated_and_kahan(ctx, ss, kahan):
    c = case("ram_fs0", seed=11, nt=301 if ss > 1 else 300, fs=1)
    say(f"[gpu] two steps per launch: store_snap {ss} kahan {kahan}")
    Rec, ko = oracle_case(c, store_snap=ss)
    k1 = engine_image(ctx, c, Rec, 1, store_snap=ss, kahan=kahan)
    k2 = engine_image(ctx, c, Rec, 2, store_snap=ss, kahan=kahan)
    np.testing.assert_array_equal(k2, k1)
    assert np.linalg.norm(k1) > 0
    if not kahan:
        np.testing.assert_array_equal(k2, ko)


def test_two_steps_shot_mix_groups_and_admission(ctx):
    """The five uneven shots of the engine tests through 2 slots (admission mid-run), in one
    and two launch chains: every image equals the one-step engine's and the oracle's."""
    from tests.test_gpu_engine import SHOTS, model, oracle_image, run_engine, shot_inputs

    vel = model(7)
    say("[gpu] two steps per launch: 5 uneven shots, 2 slots, groups 1 / 2")
    ref = {q: oracle_image(shot_inputs(vel, SHOTS[q], q)) for q in range(5)}
    for groups in (1, 2):
        imgs, _, st = run_engine(ctx, vel, SHOTS, range(5), nslots=2, groups=groups, chunk_steps=16,
                                 steps_per_launch=2)
        for q in range(5):
            np.testing.assert_array_equal(imgs[q], ref[q], err_msg=f"shot {q} groups {groups}")
        assert st["slot_steps"] == sum(2 * (s[2] - 1) for s in SHOTS)


def test_two_steps_production_windows(ctx):
    from tests.test_gpu_engine_production import NTS, line, oracle_images, run_engine, shot_job

    W, pick = line()
    m = W["model"]
    jobs = [shot_job(m, W["shots"][k], nt, k) for k, nt in zip(pick, NTS)]
    say(f"[gpu] two steps per launch: windows {[j['nx_w'] for j in jobs]} x 601, nt {NTS}")
    ref = oracle_images(jobs)
    imgs, _, _ = run_engine(ctx, m, jobs, 3, groups=2, chunk_steps=32, steps_per_launch=2)
    for j, a, b in zip(jobs, imgs, ref):
        np.testing.assert_array_equal(a, b, err_msg=f"window {j['nx_w']}")


def test_two_steps_rejects_16_bit_snapshots(ctx):
    import torch

    from psdm_amd._native import RtmError
    from psdm_amd.engine import Engine

    d_vel = torch.ones((60, 40), device="cuda:0")
    with pytest.raises(RtmError, match="FP32 snapshots"):
        Engine(ctx, d_vel, 40, 60, nslots=1, max_nx=40, max_nt=50, dPML=10, fs=0, dmodel=10.0, dt=0.001,
               snap_format="f16", store_snap=4, steps_per_launch=2)
