"""Writes the golden vectors of the RTM solver path from the reference Fortran.

Runs the reference's own solvers (solver_ac_{forward,backward}_{ram,disk}, compiled
from their unmodified sources into oracle/_ref/libpsdm_ref.so by `make -C oracle ref`)
on the seeded cases of tests/cases.py and stores inputs and outputs as .npz fixtures:
traces Rec_point, a few snapshots of Fields, the image kernel, plus the reference's
coefficient tables.  Run in the development container (where the reference is
mounted):  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import pyoracle as O  # noqa: E402
from tests.cases import CASES, case  # noqa: E402

SNAP_PICK = 6  # snapshots kept per case (evenly spaced)


def run(name):
    c = case(name)
    geo = (c["nxSou"], c["nySou"], c["nxRec"], c["nyRec"])
    tmp = tempfile.mkdtemp(prefix="psdm_golden_")
    cwd = os.getcwd()
    os.chdir(tmp)  # the _disk solvers write fort.22 and snapshot files into the cwd/folder
    try:
        kw = dict(variant=c["variant"], impl="ref")
        if c["variant"] == 1:
            kw["folder"] = tmp
        Rec, Fields = O.forward(*geo, c["S"], c["dx"], c["dt"], c["ctmp"], c["store_snap"], c["nt_snap"], c["nt"],
                                c["dPML"], c["fs"], **kw)
        Sadj = np.asfortranarray(Rec[::-1, :])  # BackSource = Data_target(nt:1:-1) (calculate_PSDM_RTM.f90:448-450)
        kernel = O.backward(*geo, Sadj, c["dx"], c["dt"], c["ctmp"], c["store_snap"], c["nt_snap"], c["nt"],
                            c["dPML"], c["fs"], Fields=Fields, **kw)
    finally:
        os.chdir(cwd)
        shutil.rmtree(tmp, ignore_errors=True)
    out = {k: np.asarray(v) for k, v in c.items() if isinstance(v, np.ndarray)}
    out.update({k: np.asarray(v) for k, v in c.items() if not isinstance(v, np.ndarray)})
    out["Rec"] = Rec
    out["kernel"] = kernel
    if Fields is not None:
        nfill = min(c["nt_snap"], (c["nt"] - 1) // c["store_snap"])
        pick = np.unique(np.linspace(0, nfill - 1, SNAP_PICK).astype(int))
        out["snap_idx"] = pick
        out["snaps"] = np.ascontiguousarray(Fields[:, :, pick])
    return out


def main():
    for name in CASES:
        d = run(name)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **d)
        print(name, "Rec", d["Rec"].shape, "kernel |.|", float(np.linalg.norm(d["kernel"])))
    # coefficient tables of the reference (PML_coeffs, finite_difference_coef)
    tabs = {}
    for dPML, dt, d in ((10, 0.001, 10.0), (40, 0.0008, 10.0), (40, 0.001, 12.5), (20, 0.0013, 12.5)):
        fA, fB, cA, cB = O.ref_pml_tables(dPML, dt, d)
        key = f"d{dPML}_{dt}_{d}"
        tabs[key + "_fA"], tabs[key + "_fB"], tabs[key + "_cA"], tabs[key + "_cB"] = fA, fB, cA, cB
    tabs["fd1"] = O.ref_fd_coef(1)
    tabs["fd2"] = O.ref_fd_coef(2)
    rng = np.random.default_rng(7)
    p = rng.standard_normal((23, 19)).astype(np.float32)
    tabs["deriv_in"] = p
    for w in range(4):
        tabs[f"deriv_{w}"] = O.ref_deriv(w, p)
    np.savez_compressed(os.path.join(HERE, "tables.npz"), **tabs)
    print("tables", len(tabs))


if __name__ == "__main__":
    main()
