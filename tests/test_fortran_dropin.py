"""The Fortran drop-in (psdm_amd/csrc/psdm_rtm_fortran.f90).

CPU: the shim links against the reference's own callers (dynamic_window.f90, compiled
from its unmodified source) and leaves none of the four solver symbols unresolved.
GPU: a Fortran caller reproducing calculate_PSDM_RTM.f90:395-466 goes through the shim
and libpsdm_rtm.so and matches the oracle bit for bit.
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as O
from tests.cases import case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"


@pytest.mark.skipif(not (os.path.isdir(REF_SRC) and os.path.exists("/opt/rocm/bin/amdflang")),
                    reason="needs the reference sources and amdflang")
def test_shim_links_against_reference_callers():
    from psdm_amd._native import LIB_PATH

    if not os.path.exists(LIB_PATH):
        from psdm_amd.build import build

        build()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "linkcheck"], check=True,
                   capture_output=True)
    so = os.path.join(ROOT, "oracle", "_ref", "linkcheck.so")
    und = subprocess.run(["nm", "-u", so], capture_output=True, text=True, check=True).stdout
    unresolved = sorted(set(w for w in und.split() if w.startswith("solver_ac")))
    # solver_ac_forward_ram_tt is the dead travel-time variant (dynamic_window_solver_TTs
    # has no caller); the four RTM solvers must all be provided by the shim.
    assert unresolved == ["solver_ac_forward_ram_tt_"], unresolved
    defined = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True).stdout
    for s in ("solver_ac_forward_ram_", "solver_ac_backward_ram_", "solver_ac_forward_disk_",
              "solver_ac_backward_disk_", "dynamic_window_solver_forward_ram_"):
        assert s in defined


def _write_input(path, c, data, taper=0, T_taper=0.0):
    ny, nx, nt = c["ny"], c["nx"], c["nt"]
    NumRec = c["nxRec"].size
    with open(path, "wb") as f:
        np.array([ny, nx, nt, NumRec, c["dPML"], c["fs"], taper, c["variant"]], np.int32).tofile(f)
        np.array([c["dx"], c["dt"], T_taper], np.float32).tofile(f)
        np.array([c["nxSou"][0], c["nySou"][0]], np.int32).tofile(f)
        c["nxRec"].astype(np.int32).tofile(f)
        c["nyRec"].astype(np.int32).tofile(f)
        c["S"][:, 0].astype(np.float32).tofile(f)
        c["ctmp"].astype(np.float32).ravel(order="F").tofile(f)
        data.astype(np.float32).ravel(order="F").tofile(f)


@pytest.mark.gpu
@pytest.mark.parametrize("name,taper", [("ram_fs0", 0), ("ram_fs1", 0), ("disk_fs1", 0), ("ram_fs0", 1),
                                        ("ram_fs1", 1)])
def test_fortran_caller_through_shim_matches_oracle(name, taper):
    """With taper_option != 0 the snapshots are tapered on the device (never copied to the
    host); the oracle tapers its own snapshots as the driver tapers Fields
    (calculate_PSDM_RTM.f90:421-427; oracle/conditioning.taper_weights, pinned against the
    reference's taper_pad_trace), and the image must match it bit for bit."""
    from oracle import conditioning as OC
    from psdm_amd.build import CALLER

    assert os.path.exists(CALLER), "tests/fortran/_build/rtm_caller is built by __graft_entry__.build()"
    c = case(name)
    T_taper = 0.05
    print(f"[gpu] fortran caller case={name} taper={taper}", flush=True)
    geo = (c["nxSou"], c["nySou"], c["nxRec"], c["nyRec"])
    rng = np.random.default_rng(3)
    data = np.asfortranarray(rng.standard_normal((c["nt"], c["nxRec"].size)).astype(np.float32) * 1e-3)
    tmp = tempfile.mkdtemp()
    try:
        _write_input(os.path.join(tmp, "in.bin"), c, data, taper, T_taper)
        subprocess.run([CALLER, "in.bin", "out.bin"], cwd=tmp, check=True, timeout=120)
        out = np.fromfile(os.path.join(tmp, "out.bin"), np.float32)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    nt, NumRec, ny, nx = c["nt"], c["nxRec"].size, c["ny"], c["nx"]
    Rec = out[:nt * NumRec].reshape((nt, NumRec), order="F")
    kernel = out[nt * NumRec:].reshape((ny, nx), order="F")
    Ro, Fo = O.forward(*geo, c["S"], c["dx"], c["dt"], c["ctmp"], 1, nt, nt, c["dPML"], c["fs"],
                       variant=c["variant"], table_mode=0)
    if taper:
        N_taper = max(1, int(np.float32(T_taper) / np.float32(c["dt"])))  # calculate_PSDM_RTM.f90:287
        w = OC.taper_weights(N_taper)
        i0 = max(1, nt - N_taper + 1)
        for n in range(N_taper):
            if i0 + n <= nt:
                Fo[:, :, i0 + n - 1] = (Fo[:, :, i0 + n - 1] * w[n]).astype(np.float32)
        assert N_taper > 1
    ko = O.backward(*geo, np.asfortranarray(data[::-1, :]), c["dx"], c["dt"], c["ctmp"], 1, nt, nt, c["dPML"],
                    c["fs"], Fields=Fo, variant=c["variant"], table_mode=0)
    np.testing.assert_array_equal(Rec, Ro)
    np.testing.assert_array_equal(kernel, ko)


@pytest.mark.gpu
def test_fortran_conditioning_binding_matches_reference():
    """The driver's conditioning block as one Fortran call (psdm_condition_gather) on the
    golden cases written from the reference's own routines."""
    from psdm_amd.build import COND_CALLER

    assert os.path.exists(COND_CALLER), "tests/fortran/_build/cond_caller is built by __graft_entry__.build()"
    z = np.load(os.path.join(ROOT, "tests", "golden", "conditioning_ref.npz"))
    for name in [k[:-4] for k in z.files if k.endswith("_raw")]:
        print(f"[gpu] fortran conditioning case={name}", flush=True)
        p = dict(zip(z["param_names"], z[name + "_params"]))
        raw = np.asfortranarray(z[name + "_raw"])
        nt_data, n = raw.shape
        nt_taper = int(p["nt_taper"])
        sf = z[name + "_nt_sf"] if name + "_nt_sf" in z.files else np.zeros(n, np.int32)
        cut = z[name + "_nt_cut"] if name + "_nt_cut" in z.files else np.zeros(n, np.int32)
        tmp = tempfile.mkdtemp()
        try:
            with open(os.path.join(tmp, "in.bin"), "wb") as f:
                np.array([nt_data, n, int(p["typef"]), int(p["nt_prop2"]), nt_taper, int(p["taper_option"]),
                          int(p["N_taper"]), int(p["N_zeros"]), int(name + "_nt_sf" in z.files),
                          int(name + "_nt_cut" in z.files)], np.int32).tofile(f)
                np.array([p["dt_data"], p["f_init"], p["f_final"], p["dt_prop"]], np.float32).tofile(f)
                sf.astype(np.int32).tofile(f)
                cut.astype(np.int32).tofile(f)
                raw.astype(np.float32).ravel(order="F").tofile(f)
            subprocess.run([COND_CALLER, "in.bin", "out.bin"], cwd=tmp, check=True, timeout=120)
            out = np.fromfile(os.path.join(tmp, "out.bin"), np.float32).reshape((nt_taper, n), order="F")
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        ref = z[name + "_out"]
        if float(p["f_init"]) > 0:
            assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-5, name
        else:
            np.testing.assert_array_equal(out, ref, err_msg=name)


@pytest.mark.gpu
def test_fortran_su_reader_binding_matches_reference_reader():
    """The driver's shot read (open_shot_mcs + get_shot_data) as one Fortran call
    (psdm_get_shot_data) on big- and little-endian files, with and without the data delay;
    against the oracle reader (pinned to the reference's get_shot_data)."""
    from oracle import su_codec as S
    from psdm_amd.build import SU_CALLER

    assert os.path.exists(SU_CALLER), "tests/fortran/_build/su_caller is built by __graft_entry__.build()"
    rng = np.random.default_rng(4)
    for big, nrec, ntmax, nt, delay, dt in ((1, 96, 1200, 1100, 0.0, 0.002), (0, 33, 700, 700, 0.05, 0.002),
                                            (1, 5, 64, 60, -0.012, 0.004)):
        print(f"[gpu] fortran su reader big={big} nrec={nrec} nt={nt} delay={delay}", flush=True)
        e = ">" if big else "<"
        buf = b"".join(rng.integers(0, 99, 60).astype(e + "i4").tobytes()
                       + rng.standard_normal(ntmax).astype(e + "f4").tobytes() for _ in range(nrec))
        tmp = tempfile.mkdtemp()
        try:
            with open(os.path.join(tmp, "shot.su"), "wb") as f:
                f.write(buf)
            subprocess.run([SU_CALLER, tmp + "/", "shot.su", str(nrec), str(ntmax), str(nt), str(big), repr(delay),
                            repr(dt), "out.bin"], cwd=tmp, check=True, timeout=120)
            o = np.fromfile(os.path.join(tmp, "out.bin"), np.float32)
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        raw = o[:nrec * nt].reshape(nrec, nt)
        dl = o[nrec * nt:].reshape(nrec, nt)
        np.testing.assert_array_equal(raw, S.decode(buf, nrec, ntmax, nt, bool(big)))
        if delay != 0.0:
            shift = int(np.float32(delay) / np.float32(dt))
            np.testing.assert_array_equal(dl, S.decode(buf, nrec, ntmax, nt, bool(big), shift))
        else:
            assert (dl == -1.0).all()  # untouched, as the reference's caller ignores it then


@pytest.mark.gpu
def test_fortran_travel_time_binding_matches_reference_eikonal():
    """The Kirchhoff driver's travel-time calls (dynamic_window_tt_forward / _backward) as
    psdm_window_tt_forward / _backward from a Fortran caller: the shot window, then the
    source's and every receiver's travel times, bit for bit the oracle's solve_eikonal_FSM
    (pinned to the reference's)."""
    from oracle import kirchhoff as K
    from psdm_amd.build import TT_CALLER

    assert os.path.exists(TT_CALLER), "tests/fortran/_build/tt_caller is built by __graft_entry__.build()"
    rng = np.random.default_rng(8)
    ny, nx, dd, dd_space = 45, 200, 10.0, 150.0
    vel = np.asfortranarray((1500 + 1500 * rng.random((ny, nx))).astype(np.float32))
    nxSou, nySou = 90, 2
    nxRec = np.arange(95, 95 + 12 * 5, 5, dtype=np.int32)
    nyRec = np.full(nxRec.size, 3, np.int32)
    print(f"[gpu] fortran travel times {ny}x{nx}, {nxRec.size} receivers", flush=True)
    tmp = tempfile.mkdtemp()
    try:
        with open(os.path.join(tmp, "in.bin"), "wb") as f:
            np.array([ny, nx, nxRec.size, nxSou, nySou], np.int32).tofile(f)
            nxRec.tofile(f)
            nyRec.tofile(f)
            np.array([dd, dd_space], np.float32).tofile(f)
            vel.ravel(order="F").tofile(f)
        subprocess.run([TT_CALLER, "in.bin", "out.bin"], cwd=tmp, check=True, timeout=120)
        raw = open(os.path.join(tmp, "out.bin"), "rb").read()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    x_i, nx_w = np.frombuffer(raw[:8], np.int32)
    tt = np.frombuffer(raw[8:], np.float32)
    assert nx_w == min(nx, nxRec.max() + 16) - max(1, nxSou - 16 + 1) + 1 and x_i == nxSou - 15
    cols = np.ascontiguousarray(vel[:, x_i - 1:x_i - 1 + nx_w].T)  # (nx_w, ny)
    ts = tt[:ny * nx_w].reshape(nx_w, ny)
    np.testing.assert_array_equal(ts.view(np.uint32), K.eikonal(cols, dd, dd, nxSou - x_i + 1, nySou).view(np.uint32))
    tr = tt[ny * nx_w:].reshape(nxRec.size, nx_w, ny)
    for q in (0, 5, nxRec.size - 1):
        want = K.eikonal(cols, dd, dd, int(nxRec[q]) - x_i + 1, int(nyRec[q]))
        np.testing.assert_array_equal(tr[q].view(np.uint32), want.view(np.uint32))
