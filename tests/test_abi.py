"""The C ABI library: it builds, loads, exports every entry point declared in
include/psdm_rtm.h, and its host-side error paths work without a GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.conftest import gpu_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "psdm_rtm.h")


def _lib_path():
    from psdm_amd import _native

    if not os.path.exists(_native.LIB_PATH):
        from psdm_amd.build import build

        build()
    return _native.LIB_PATH


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(psdm_rtm_\w+)\s*\(", text)))


def test_header_declares_the_solver_seam():
    syms = declared_symbols()
    for name in ("psdm_rtm_forward_ram", "psdm_rtm_backward_ram", "psdm_rtm_forward_disk",
                 "psdm_rtm_backward_disk", "psdm_rtm_create", "psdm_rtm_destroy", "psdm_rtm_last_error"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    path = _lib_path()
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (psdm_rtm_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_every_symbol():
    from psdm_amd import _native

    _lib_path()
    L = _native.lib()
    for s in declared_symbols():
        assert s in _native.SIGNATURES, s
        getattr(L, s)


def _struct_fields(name):
    """Member names, in order, of `typedef struct { ... } name;` in the header."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct \{([^}]*)\}\s*" + name + r"\s*;", text).group(1)
    names = []
    for decl in body.split(";"):
        decl = re.sub(r"^\s*(const\s+)?(long\s+long|\w+)", "", decl)  # drop the type
        names += [d.strip(" *\n") for d in decl.split(",") if d.strip()]
    return names


def test_python_structs_mirror_the_header():
    from psdm_amd.engine import EngineCfg, Job

    assert [f[0] for f in EngineCfg._fields_] == _struct_fields("psdm_rtm_engine_cfg")
    assert [f[0] for f in Job._fields_] == _struct_fields("psdm_rtm_job")
    from psdm_amd.conditioning import Cond

    assert [f[0] for f in Cond._fields_] == _struct_fields("psdm_rtm_cond")


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_context_without_device_fails_loudly():
    from psdm_amd._native import Context, RtmError

    _lib_path()
    with pytest.raises(RtmError):
        Context(0)


def test_dynamic_window_matches_reference_rule():
    """dynamic_window.f90:15-50 on the 401 x 201 @ 12.5 m single-shot geometry: source
    at x = 500 m (column 41), 240 channels from 100 m offset at 12.5 m toward +x."""
    from psdm_amd.solvers import dynamic_window

    dx = 12.5
    nxS = 1 + int(np.floor(500.0 / dx))
    pos = 500.0 + (100.0 + np.arange(240) * 12.5)  # dir_boat = -1: pos_rec = pos_shot + offsets
    nxR = 1 + np.floor(pos / dx).astype(int)
    nx_w, x_i, x_f = dynamic_window([nxS], nxR, dx, 401, 500.0)
    assert (nxS, nxR.min(), nxR.max()) == (41, 49, 288)
    ddx = 1 + int(500.0 // 12.5)
    assert (x_i, x_f) == (max(41 - ddx + 1, 1), min(288 + ddx, 401))
    assert nx_w == x_f - x_i + 1 == 329
    # clipping at both model edges
    nx_w2, xi2, xf2 = dynamic_window([3], [5, 9], 10.0, 15, 100.0)
    assert (xi2, xf2, nx_w2) == (1, 15, 15)


@pytest.mark.skipif(not __import__("oracle.pyoracle", fromlist=["x"]).ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("f0,dt,t0,nt", [(15.0, 0.001, 0.0, 3000), (15.0, 0.001, 0.08, 900), (8.0, 0.001, 0.1875, 4000),
                                         (25.0, 0.0008, 0.0, 6000), (10.0, 0.0005, 0.12, 2500),
                                         (8.0, 0.001, 0.1875, 12000), (15.0, 0.0008, 0.08, 6000)])
def test_ricker_mcs_vs_reference(f0, dt, t0, nt):
    """ricker_mcs against the reference's own Ricker_mcs (oracle/_ref) with the module
    inputs freq_ricker, delay_ricker, nt_prop, dt_prop set: bit for bit, zero tail from
    t >= 10/f0 included."""
    from oracle.pyoracle import ref_ricker_mcs
    from psdm_amd.solvers import ricker_mcs

    got, ref = ricker_mcs(nt, dt, f0, t0), ref_ricker_mcs(nt, dt, f0, t0)
    np.testing.assert_array_equal(got, ref)
    # a wrong factor in (1 - 2a) exp(-a), e.g. (1 - a) exp(-a), moves the side lobes far off
    t = np.arange(nt, dtype=np.float32) * np.float32(dt)
    a = (np.pi * f0 * (t - t0)) ** 2
    bad = ((1 - a) * np.exp(-a)).astype(np.float32)
    bad[t >= 10.0 / f0] = 0
    assert np.max(np.abs(bad - ref)) > 0.1


@pytest.mark.skipif(not __import__("oracle.pyoracle", fromlist=["x"]).ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("f0,dt,nt", [(8.0, 0.001, 12000), (7.0, 0.0008, 3000)])
def test_ricker_was_vs_reference_expression(f0, dt, nt):
    """Ricker_was is Ricker_mcs's expression with t0 = 1.5/f0 (get_was_data.f90:484): the
    reference's Ricker_mcs run with delay_ricker = FP32(1.5/f0) gives it bit for bit."""
    from oracle.pyoracle import ref_ricker_mcs
    from psdm_amd.solvers import ricker_was

    np.testing.assert_array_equal(ricker_was(nt, dt, f0), ref_ricker_mcs(nt, dt, f0, np.float32(1.5) / np.float32(f0)))


def test_ricker_mcs():
    from psdm_amd.solvers import ricker_mcs

    s = ricker_mcs(400, 0.001, 15.0, 0.1)
    assert s.dtype == np.float32
    assert s[100] == pytest.approx(1.0)  # peak at t0
    assert np.argmax(s) == 100
    assert np.all(s[np.arange(400) * 0.001 >= 10.0 / 15.0] == 0)
    # symmetric about t0 to FP32 rounding
    np.testing.assert_allclose(s[100 - 30:100], s[101:131][::-1], rtol=1e-5, atol=1e-7)
