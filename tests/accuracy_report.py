"""Accuracy of the batched engine's snapshot formats against the FP64 oracle on full-size
shots of the uneven 256-shot line (601-deep model @ 10 m, dt 0.8 ms, dPML 40).

For a sample of shots (by default the three narrowest windows among the nt = 8000
shots, so all three reflectors are imaged) the observed gather is modelled by the FP32
oracle on the true and on the water-only model (direct wave removed) and migrated by the
FP64 oracle with snapshots every 4th step (backward step k paired with the snapshot of
forward step nt + 2 - k, the engine's decimated bookkeeping).  The same gathers are
migrated on the GPU by the engine with FP32, FP16 and BF16 snapshots every 4th step and
Kahan-compensated image accumulation, each with the reference's second-derivative
weights and with the opt-in zero-sum weights (`<fmt>_zero_sum`).  Reported per format and shot: image rel-L2 vs
FP64 and the RMS amplitude error along each known reflector (peak |image| of the FP64
image within +-5 samples of the reflector depth, compared at the same sample).

The oracle runs in a process pool started before anything touches the GPU; as it takes
minutes per full-size shot, it can run apart from the GPU part: `--oracle-out DIR` writes
its results (no GPU needed), `--ref DIR` reads them and runs only the GPU part.  Prints
one JSON object.  Usage (repo root):
    python -m tests.accuracy_report [--shots 3] [--nt 8000] [--oracle-out DIR | --ref DIR]
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def shot_payload(W, k, nt):
    from psdm_amd.solvers import ricker_mcs

    m, sh = W["model"], W["shots"][k]
    return dict(k=k, nt=nt, ss=4, dx=m.dx, dt=W["dt"], dPML=W["dPML"], nxSou=sh.nxSou - sh.x_i + 1, nySou=sh.nySou,
                nxRec=sh.nxRec - sh.x_i + 1, nyRec=sh.nyRec, S=ricker_mcs(nt, W["dt"], W["f0"], 1.2 / W["f0"]),
                ctmp=np.asfortranarray(m.migration[:, sh.x_i - 1:sh.x_f]),
                ctrue=np.asfortranarray(m.true[:, sh.x_i - 1:sh.x_f]),
                cwat=np.asfortranarray(m.water[:, sh.x_i - 1:sh.x_f]))


def oracle_worker(p):
    """Models the gather (FP32 oracle, direct wave removed) and migrates it in FP64."""
    from oracle import pyoracle as O

    nt, ss = p["nt"], p["ss"]
    geo = ([p["nxSou"]], [p["nySou"]], p["nxRec"], p["nyRec"])
    S = p["S"].reshape(nt, 1)
    t0 = time.time()
    d_true, _ = O.forward(*geo, S, p["dx"], p["dt"], p["ctrue"], 1, 0, nt, p["dPML"], 0)
    d_wat, _ = O.forward(*geo, S, p["dx"], p["dt"], p["cwat"], 1, 0, nt, p["dPML"], 0)
    data = np.asfortranarray(d_true - d_wat)
    nsnap = (nt - 1) // ss
    _, F = O.forward(*geo, S, p["dx"], p["dt"], p["ctmp"], ss, nsnap, nt, p["dPML"], 0, precision="f64")
    k64 = O.backward(*geo, np.asfortranarray(data[::-1, :]), p["dx"], p["dt"], p["ctmp"], ss, nsnap, nt, p["dPML"], 0,
                     Fields=F, variant=2, precision="f64")
    return dict(k=p["k"], data=data, image64=k64, seconds=time.time() - t0)


def reflector_amplitude_errors(img, ref, reflectors, x_i, dx, halfwin=5):
    ny, nx = ref.shape
    out = []
    for (z0, dip, dv) in reflectors:
        a, b = [], []
        for i in range(nx):
            x = (x_i - 1 + i) * dx
            jz = int(round((z0 + np.tan(np.radians(dip)) * x) / dx))
            lo, hi = max(jz - halfwin, 0), min(jz + halfwin + 1, ny)
            if hi - lo < 2 * halfwin + 1:
                continue
            r = np.abs(ref[lo:hi, i])
            q = int(np.argmax(r))
            a.append(abs(float(img[lo + q, i])))
            b.append(float(r[q]))
        a, b = np.asarray(a), np.asarray(b)
        err = float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2))) if b.size and np.any(b) else None
        out.append({"z0_m": round(float(z0), 1), "dip_deg": dip, "dv": dv, "columns": int(b.size), "rms_rel_err": err})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--shots", type=int, default=3)
    ap.add_argument("--nt", type=int, default=8000)
    ap.add_argument("--oracle-out", default=None, help="compute the oracle part only and save it here")
    ap.add_argument("--ref", default=None, help="read the oracle part from here and run the GPU part")
    args = ap.parse_args()

    from psdm_amd.workloads import uneven_line

    W = uneven_line(256, seed=0)
    shots = W["shots"]
    cand = sorted([k for k in range(len(shots)) if W["nt"][k] == args.nt], key=lambda k: shots[k].nx_w)
    idx = cand[:args.shots]
    payloads = [shot_payload(W, k, args.nt) for k in idx]
    if args.ref:
        ref = {}
        for k in idx:
            z = np.load(os.path.join(args.ref, f"oracle_shot{k}_nt{args.nt}.npz"))
            ref[k] = dict(k=k, data=np.asfortranarray(z["data"]), image64=np.asfortranarray(z["image64"]))
    else:
        log(f"[accuracy] FP64 oracle on shots {idx} (windows {[shots[k].nx_w for k in idx]}, nt {args.nt})")
        pool = mp.get_context("spawn").Pool(len(payloads))  # before any GPU use in this process
        ref = {}
        for r in pool.imap_unordered(oracle_worker, payloads):
            ref[r["k"]] = r
            log(f"[accuracy] oracle shot {r['k']} done in {r['seconds']:.0f} s")
            if args.oracle_out:
                os.makedirs(args.oracle_out, exist_ok=True)
                np.savez(os.path.join(args.oracle_out, f"oracle_shot{r['k']}_nt{args.nt}.npz"), data=r["data"],
                         image64=r["image64"])
        pool.close()
        pool.join()
        if args.oracle_out:
            return

    import torch

    from psdm_amd._native import Context
    from psdm_amd.engine import Engine
    from tests.cases import rel_l2

    dev = torch.device("cuda:0")
    ctx = Context(0)
    m = W["model"]
    d_mig = torch.from_numpy(np.ascontiguousarray(m.migration.T)).to(dev)
    report = {"workload": "uneven line sample, 601-deep @10 m, dt 0.8 ms, dPML 40, snapshots every 4th step",
              "shots": [{"index": k, "nx_w": shots[k].nx_w, "nt": args.nt} for k in idx], "forms": {}}
    for fmt, weights in [(f, w) for w in ("reference", "zero_sum") for f in ("f32", "f16", "bf16")]:
        eng = Engine(ctx, d_mig, m.nz, m.nx, nslots=len(idx), max_nx=max(shots[k].nx_w for k in idx),
                     max_nt=args.nt, dPML=W["dPML"], fs=W["fs"], dmodel=m.dx, dt=W["dt"], tile_x=32,
                     snap_format=fmt, store_snap=4, kahan=True, weights=weights)
        imgs, keep = [], []
        for k in idx:
            sh, p = shots[k], payloads[idx.index(k)]
            d_w = torch.from_numpy(p["S"]).to(dev)
            d_d = torch.from_numpy(np.ascontiguousarray(ref[k]["data"].T)).to(dev)
            d_i = torch.zeros((sh.nx_w, m.nz), device=dev)
            eng.submit(sh.x_i, sh.nx_w, args.nt, p["nxSou"], p["nySou"], p["nxRec"], p["nyRec"], d_w, d_d, d_i)
            imgs.append(d_i)
            keep += [d_w, d_d]
        eng.drain()
        eng.close()
        res = []
        for k, d_i in zip(idx, imgs):
            img = d_i.cpu().numpy().T.copy(order="F")
            res.append({"index": k, "image_rel_l2_vs_fp64": rel_l2(img, ref[k]["image64"]),
                        "reflectors": reflector_amplitude_errors(img, ref[k]["image64"], m.reflectors, shots[k].x_i,
                                                                 m.dx)})
        name = fmt if weights == "reference" else f"{fmt}_zero_sum"
        report["forms"][name] = res
        log(f"[accuracy] {name}: image rel-L2 {[round(r['image_rel_l2_vs_fp64'], 7) for r in res]}")
    print(json.dumps(report), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
