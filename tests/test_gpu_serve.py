"""GPU test of the served form (psdm_amd.serve over two device-resident engines).

A stream of requests with Poisson arrivals is served by two engines (two velocity
models: FP32 snapshots every step, and FP16 snapshots every 4th step with Kahan image
accumulation).  Requests overlap on the device; each request's image must be the same,
bit for bit, as the image of that request migrated alone, and the FP32 ones must equal
the oracle's one-shot solver image.
"""
import sys

import numpy as np
import pytest

from tests.test_gpu_engine import DPML, DT, DX, NXP, NY, SHOTS, model, oracle_image, shot_inputs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


ENGINE_KW = {"a": dict(snap_format="f32", store_snap=1, kahan=False),
             "b": dict(snap_format="f16", store_snap=4, kahan=True)}


def make_engines(ctx, vels, nslots):
    import torch

    from psdm_amd.engine import Engine

    dev = torch.device("cuda:0")
    engines, keep = {}, []
    for name, vel in vels.items():
        d_vel = torch.from_numpy(np.ascontiguousarray(vel.T)).to(dev)
        keep.append(d_vel)
        engines[name] = Engine(ctx, d_vel, NY, NXP, nslots=nslots, max_nx=max(s[1] for s in SHOTS),
                               max_nt=max(s[2] for s in SHOTS), dPML=DPML, fs=0, dmodel=DX, dt=DT, chunk_steps=16,
                               **ENGINE_KW[name])
    return engines, keep


def request_buffers(vel, q):
    import torch

    dev = torch.device("cuda:0")
    si = shot_inputs(vel, SHOTS[q % len(SHOTS)], q)
    job = dict(x_i=si["x_i"], nx_w=si["nx_w"], nt=si["nt"], nxSou=si["nxSou"], nySou=si["nySou"],
               nxRec=si["nxRec"], nyRec=si["nyRec"], d_wavelet=torch.from_numpy(si["S"]).to(dev),
               d_data=torch.from_numpy(np.ascontiguousarray(si["data"].T)).to(dev),
               d_image_out=torch.zeros((si["nx_w"], NY), device=dev))
    return si, job


def test_served_stream_results_equal_alone_and_oracle(ctx):
    import torch

    from psdm_amd.serve import Request, Server, poisson_arrivals

    vels = {"a": model(6), "b": model(7)}
    n = 10
    arrivals = poisson_arrivals(400.0, n, seed=1)
    route = ["a" if q % 3 else "b" for q in range(n)]
    engines, keep = make_engines(ctx, vels, nslots=3)
    reqs, infos = [], []
    for q in range(n):
        si, job = request_buffers(vels[route[q]], q)
        reqs.append(Request(q, float(arrivals[q]), route[q], job, kind=route[q]))
        infos.append(si)
    torch.cuda.synchronize()
    say(f"[gpu] serve: {n} requests over 2 engines, arrivals up to {arrivals[-1] * 1e3:.1f} ms")
    out = Server(engines).run(reqs)
    torch.cuda.synchronize()
    served = [r.job["d_image_out"].cpu().numpy().T.copy(order="F") for r in reqs]
    st = {k: e.stats() for k, e in engines.items()}
    for e in engines.values():
        e.close()
    assert out["n"] == n and out["latency"]["n"] == n
    assert all(r["latency"] >= 0 for r in out["requests"])
    assert all(s["hbm_used"] == 0 for s in st.values())
    # overlap: some request was admitted while another was still on the device
    spans = sorted((r["submitted"], r["done"]) for r in out["requests"])
    assert any(spans[q + 1][0] < spans[q][1] for q in range(n - 1))

    say("[gpu] serve: each request alone")
    for q in range(n):
        eng1, keep1 = make_engines(ctx, {route[q]: vels[route[q]]}, nslots=1)
        _, job = request_buffers(vels[route[q]], q)
        eng1[route[q]].submit(**job)
        eng1[route[q]].drain()
        alone = job["d_image_out"].cpu().numpy().T.copy(order="F")
        eng1[route[q]].close()
        np.testing.assert_array_equal(served[q], alone)
        assert np.linalg.norm(alone) > 0
        if route[q] == "a" and q < 4:
            np.testing.assert_array_equal(served[q], oracle_image(infos[q]))
