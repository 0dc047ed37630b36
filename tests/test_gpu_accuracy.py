"""GPU accuracy against the FP64 oracle on the single-shot 401 x 201 @ 12.5 m workload
(window 201 x 329, dPML 40, 600 steps): the plain solver path and the batched engine
(FP32 snapshots every step) are compared with the FP64 oracle image and traces kept in
tests/golden/accuracy_single401_fp64.npz.  Bar: rel-L2 <= 1e-5 (FP32 arithmetic)."""
import os

import numpy as np
import pytest

from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def inputs():
    from psdm_amd import solvers as S
    from psdm_amd.workloads import line_workload

    fx = np.load(os.path.join(HERE, "golden", "accuracy_single401_fp64.npz"))
    W = line_workload("single401")
    m, sh = W["model"], W["shots"][0]
    nt = int(fx["nt"])
    return dict(fx=fx, W=W, m=m, sh=sh, nt=nt, dt=W["dt"], dPML=W["dPML"],
                ctmp=np.asfortranarray(m.migration[:, sh.x_i - 1:sh.x_f]),
                ctrue=np.asfortranarray(m.true[:, sh.x_i - 1:sh.x_f]),
                src=S.ricker_mcs(nt, W["dt"], W["f0"], 1.2 / W["f0"]).reshape(nt, 1),
                nxS=[sh.nxSou - sh.x_i + 1], nyS=[sh.nySou], nxR=sh.nxRec - sh.x_i + 1, nyR=sh.nyRec)


def test_plain_and_engine_within_1e5_of_fp64(ctx):
    import torch

    from psdm_amd import solvers as S
    from psdm_amd.engine import Engine

    I = inputs()
    fx, m, nt, dt = I["fx"], I["m"], I["nt"], I["dt"]
    ny, nx = I["ctmp"].shape
    NumRec = I["nxR"].size
    par = S.ParFile(dPML=I["dPML"], fs=0)
    geo = (0, 1, NumRec, I["nxS"], I["nyS"], I["nxR"], I["nyR"])
    print(f"[gpu] accuracy single401 {ny}x{nx} nt={nt}", flush=True)
    d = np.zeros((nt, NumRec), np.float32, order="F")
    F = np.zeros((ny, nx, nt), np.float32, order="F")
    S.solver_ac_forward_ram(*geo, I["src"], m.dx, dt, ny, nx, I["ctrue"], 1, nt, nt, F, d, par=par,
                            host_fields=False, ctx=ctx)
    # the gather the fixture was built from
    assert float(np.sum(d.astype(np.float64) ** 2)) == pytest.approx(float(fx["data_sumsq"]), rel=1e-12)
    np.testing.assert_array_equal(d.reshape(-1, order="F")[fx["data_pick"]], fx["data_sample"])
    Rec = np.zeros((nt, NumRec), np.float32, order="F")
    k = np.zeros((ny, nx), np.float32, order="F")
    S.solver_ac_forward_ram(*geo, I["src"], m.dx, dt, ny, nx, I["ctmp"], 1, nt, nt, F, Rec, par=par,
                            host_fields=False, ctx=ctx)
    S.solver_ac_backward_ram(*geo, np.asfortranarray(d[::-1, :]), m.dx, dt, ny, nx, I["ctmp"], 1, nt, nt, F, k,
                             par=par, host_fields=False, ctx=ctx)
    e_img, e_rec = rel_l2(k, fx["image_fp64"]), rel_l2(Rec, fx["traces_fp64"])
    print(f"  plain: image {e_img:.3e}, traces {e_rec:.3e} vs FP64", flush=True)
    assert e_img <= 1e-5 and e_rec <= 1e-5

    dev = torch.device("cuda:0")
    sh = I["sh"]
    d_vel = torch.from_numpy(np.ascontiguousarray(m.migration.T)).to(dev)
    eng = Engine(ctx, d_vel, m.nz, m.nx, nslots=1, max_nx=sh.nx_w, max_nt=nt, dPML=I["dPML"], fs=0, dmodel=m.dx,
                 dt=dt, tile_x=32)
    d_img = torch.zeros((sh.nx_w, m.nz), device=dev)
    d_w = torch.from_numpy(np.ascontiguousarray(I["src"][:, 0])).to(dev)
    d_d = torch.from_numpy(np.ascontiguousarray(d.T)).to(dev)
    eng.submit(sh.x_i, sh.nx_w, nt, I["nxS"][0], I["nyS"][0], I["nxR"], I["nyR"], d_w, d_d, d_img)
    eng.drain()
    ke = d_img.cpu().numpy().T.copy(order="F")
    eng.close()
    np.testing.assert_array_equal(ke, k)  # the engine's FP32 form is the plain solver, bit for bit
