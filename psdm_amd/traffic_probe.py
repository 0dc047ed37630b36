"""Small fixed workloads to run under `rocprofv3 --pmc` (HBM traffic and instruction
counters per dispatch).

    python -m psdm_amd.traffic_probe copy    # 256 MiB device-to-device copy, 4 times
    python -m psdm_amd.traffic_probe plain   # one 601 x 715 (dPML 40) shot, 200 forward + 200 backward steps
    python -m psdm_amd.traffic_probe batched # 16 such shots resident in the engine (two slot groups of 8,
                                             # two concurrent dispatches per step), forward + backward

`copy` calibrates the counters: FETCH_SIZE / WRITE_SIZE of a copy of known size.
"""
from __future__ import annotations

import sys

import numpy as np


def main(what: str):
    import torch

    from psdm_amd._native import Context

    dev = torch.device("cuda:0")
    if what == "copy":
        a = torch.ones(64 * 1024 * 1024, device=dev)  # 256 MiB
        b = torch.empty_like(a)
        for _ in range(4):
            b.copy_(a)
        torch.cuda.synchronize()
        print("copy: 4 x 256 MiB read + 256 MiB written", flush=True)
        return
    from psdm_amd.solvers import DeviceShot, backward_device, forward_device, ricker_mcs
    from psdm_amd.workloads import line_workload

    W = line_workload("mcs2001")
    m, sh = W["model"], W["shots"][0]
    nt = 200
    ctx = Context(0)
    d_mig = torch.from_numpy(np.ascontiguousarray(m.migration.T)).to(dev)
    src = torch.from_numpy(ricker_mcs(nt, W["dt"], W["f0"], 1.2 / W["f0"]).reshape(1, nt)).to(dev)
    nrec = sh.nxRec.size
    rng = np.random.default_rng(0)
    data = torch.from_numpy((1e-3 * rng.standard_normal((nrec, nt))).astype(np.float32)).to(dev)
    geo = DeviceShot(m.nz, sh.nx_w, nt, W["dPML"], W["fs"], m.dx, W["dt"], [sh.nxSou - sh.x_i + 1], [sh.nySou],
                     sh.nxRec - sh.x_i + 1, sh.nyRec)
    win = d_mig[sh.x_i - 1:sh.x_f]
    torch.cuda.synchronize()
    if what == "plain":
        rec = torch.empty((nrec, nt), device=dev)
        img = torch.zeros((sh.nx_w, m.nz), device=dev)
        forward_device(ctx, geo, win, src, rec, key=1)
        backward_device(ctx, geo, win, data, img, key=1)
        ctx.sync()
        print(f"plain: {2 * (nt - 1)} step kernels on a {m.nz}x{sh.nx_w} window", flush=True)
    elif what == "batched":
        from psdm_amd.engine import Engine

        eng = Engine(ctx, d_mig, m.nz, m.nx, nslots=16, max_nx=sh.nx_w, max_nt=nt, dPML=W["dPML"], fs=W["fs"],
                     dmodel=m.dx, dt=W["dt"], tile_x=16, stream_w=4, groups=2, chunk_steps=64)
        for _ in range(16):
            eng.submit(sh.x_i, sh.nx_w, nt, sh.nxSou - sh.x_i + 1, sh.nySou, sh.nxRec - sh.x_i + 1, sh.nyRec, src[0],
                       data)
        eng.drain()
        ctx.sync()
        eng.close()
        print(f"batched: 16 slots x {2 * (nt - 1)} steps (forward + backward)", flush=True)
    ctx.close()


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "plain")
