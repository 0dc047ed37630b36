"""Writes the FP64-oracle fixtures the benchmark states its accuracy against.

* ``accuracy_single401_fp64.npz`` -- the single-shot 401 x 201 @ 12.5 m workload
  (migration window 201 x 329), cut to 600 steps.
* ``accuracy_mcs2001_fp64.npz`` -- shot 0 of the 2001 x 601 @ 10 m line (window 601 x 715,
  dt 0.8 ms) at its full record length, nt 6000.  The traces are kept for every 10th
  channel (the fixture stays small); the image whole.

In both, the observed gather is the FP32 forward modelling on the true model (the oracle
with FP64-built coefficient tables, which the GPU solver reproduces bit for bit); for
mcs2001 the water-only modelling is subtracted (direct wave removed), as the benchmark
does.  The fixture keeps the gather's sum of squares and a sample so a consumer can
confirm it modelled the same gather.  The wavelet is Ricker_mcs with the parfile's
delay_ricker set to 1.2/f0 (the wavelet starts at rest).  The FP64 oracle then models the
traces and migrates the gather on the migration model.

Run in the development container (minutes for mcs2001; about 30 GB of host memory for
its FP64 snapshots):
    python tests/golden/make_accuracy_fixture.py [single401] [mcs2001]
"""
from __future__ import annotations

import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as O  # noqa: E402
from psdm_amd.solvers import ricker_mcs  # noqa: E402
from psdm_amd.workloads import line_workload  # noqa: E402

CASES = {
    # name: (workload, shot, nt or None for the workload's own, direct wave removed, trace stride)
    "single401": ("single401", 0, 600, False, 1),
    "mcs2001": ("mcs2001", 0, None, True, 10),
}


def setup(name):
    wl, k, nt, dw, stride = CASES[name]
    W = line_workload(wl)
    m, sh = W["model"], W["shots"][k]
    nt = nt or W["nt"]
    src = ricker_mcs(nt, W["dt"], W["f0"], 1.2 / W["f0"]).reshape(nt, 1)
    geo = ([sh.nxSou - sh.x_i + 1], [sh.nySou], sh.nxRec - sh.x_i + 1, sh.nyRec)
    cols = lambda a: np.asfortranarray(a[:, sh.x_i - 1:sh.x_f])  # noqa: E731
    return dict(W=W, m=m, sh=sh, nt=nt, src=src, geo=geo, dw=dw, stride=stride, ctmp=cols(m.migration),
                ctrue=cols(m.true), cwat=cols(m.water) if dw else None)


def _model32(args):
    name, which = args
    s = setup(name)
    d, _ = O.forward(*s["geo"], s["src"], s["m"].dx, s["W"]["dt"], s[which], 1, 0, s["nt"], s["W"]["dPML"], 0)
    return which, d


def observed(name, pool=None):
    """The observed gather (nt, NumRec) FP32, as the benchmark models it on the device."""
    s = setup(name)
    jobs = [(name, "ctrue")] + ([(name, "cwat")] if s["dw"] else [])
    res = dict(pool.map(_model32, jobs) if pool else map(_model32, jobs))
    d = res["ctrue"] - res["cwat"] if s["dw"] else res["ctrue"]
    return np.asfortranarray(d.astype(np.float32))


def fp64(name, d):
    """(image, traces) of the FP64 oracle for the case's observed gather d."""
    s = setup(name)
    nt, dt, dPML, dx = s["nt"], s["W"]["dt"], s["W"]["dPML"], s["m"].dx
    R64, F64 = O.forward(*s["geo"], s["src"], dx, dt, s["ctmp"], 1, nt, nt, dPML, 0, precision="f64")
    k64 = O.backward(*s["geo"], np.asfortranarray(d[::-1, :]), dx, dt, s["ctmp"], 1, nt, nt, dPML, 0, Fields=F64,
                     precision="f64")
    return k64, R64


def write(name, pool=None):
    s = setup(name)
    d = observed(name, pool)
    k64, R64 = fp64(name, d)
    chans = np.arange(0, R64.shape[1], s["stride"])
    rng = np.random.default_rng(0)
    pick = rng.integers(0, d.size, 64)
    path = os.path.join(HERE, f"accuracy_{name}_fp64.npz")
    np.savez_compressed(path, nt=s["nt"], x_i=s["sh"].x_i, nx_w=s["sh"].nx_w, image_fp64=k64.astype(np.float32),
                        traces_fp64=R64[:, chans].astype(np.float32), trace_channels=chans,
                        direct_wave_removed=s["dw"], data_sumsq=np.sum(d.astype(np.float64) ** 2), data_pick=pick,
                        data_sample=d.reshape(-1, order="F")[pick])
    print("wrote", os.path.basename(path), k64.shape, R64[:, chans].shape, flush=True)


def main(names):
    with mp.get_context("spawn").Pool(2) as pool:
        for n in names:
            write(n, pool)


if __name__ == "__main__":
    main(sys.argv[1:] or list(CASES))
