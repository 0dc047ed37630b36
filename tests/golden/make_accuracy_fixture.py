"""Writes tests/golden/accuracy_single401_fp64.npz: the FP64 oracle's image of the
single-shot 401 x 201 @ 12.5 m workload (migration window 201 x 329, 600 steps), the
reference against which the benchmark states the GPU image's error.

The observed gather is the FP32 forward modelling on the true model (the oracle with
FP64-built coefficient tables, which the GPU solver reproduces bit for bit); the fixture
keeps its sum of squares and a sample so a consumer can confirm it modelled the same
gather.  Run in the development container:  python tests/golden/make_accuracy_fixture.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as O  # noqa: E402
from psdm_amd.solvers import ricker_mcs  # noqa: E402
from psdm_amd.workloads import line_workload  # noqa: E402

NT = 600


def main():
    W = line_workload("single401")
    m, sh = W["model"], W["shots"][0]
    dt, dPML = W["dt"], W["dPML"]
    ctmp = np.asfortranarray(m.migration[:, sh.x_i - 1:sh.x_f])
    ctrue = np.asfortranarray(m.true[:, sh.x_i - 1:sh.x_f])
    src = ricker_mcs(NT, dt, W["f0"], 1.2 / W["f0"]).reshape(NT, 1)
    geo = ([sh.nxSou - sh.x_i + 1], [sh.nySou], sh.nxRec - sh.x_i + 1, sh.nyRec)
    d, _ = O.forward(*geo, src, m.dx, dt, ctrue, 1, NT, NT, dPML, 0)
    Sadj = np.asfortranarray(d[::-1, :])
    R64, F64 = O.forward(*geo, src, m.dx, dt, ctmp, 1, NT, NT, dPML, 0, precision="f64")
    k64 = O.backward(*geo, Sadj, m.dx, dt, ctmp, 1, NT, NT, dPML, 0, Fields=F64, precision="f64")
    rng = np.random.default_rng(0)
    pick = rng.integers(0, d.size, 64)
    np.savez_compressed(os.path.join(HERE, "accuracy_single401_fp64.npz"), nt=NT, x_i=sh.x_i, nx_w=sh.nx_w,
                        image_fp64=k64.astype(np.float32), traces_fp64=R64.astype(np.float32),
                        data_sumsq=np.sum(d.astype(np.float64) ** 2), data_pick=pick,
                        data_sample=d.reshape(-1, order="F")[pick])
    print("wrote accuracy_single401_fp64.npz", k64.shape, R64.shape)


if __name__ == "__main__":
    main()
