"""Builds libpsdm_rtm.so in-tree for gfx950 with hipcc (no JIT, no torch extension).

`python -m psdm_amd.build` or `__graft_entry__.build()`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libpsdm_rtm.so")

SOURCES = ["rtm_kernels.hip", "rtm_tblock.hip", "rtm_api.hip", "rtm_engine.hip", "rtm_cond.hip", "rtm_post.hip", "rtm_su.hip", "rtm_kirch.hip"]

# -ffp-contract=off: no fused multiply-add, so each product and sum rounds separately
#   in the reference's order (the FP32 parity contract).
# -packed-fp32-ops off: plain scalar FP32 VALU ops only.
# -Werror=shorten-64-to-32: a 64-bit value (an address) silently narrowed to 32 bits is a
#   build error (the lane intrinsics, e.g. readfirstlane, take and return int).
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Werror=shorten-64-to-32",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libpsdm_rtm.so)")


def build(verbose: bool = False) -> str:
    objs = []
    cc = hipcc()
    bdir = os.path.join(PKG, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(bdir, src.replace(".hip", ".o"))
        cmd = [cc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


FLANG = "/opt/rocm/bin/amdflang"
FORTRAN_SHIM = os.path.join(CSRC, "psdm_rtm_fortran.f90")
CALLER_DIR = os.path.join(ROOT, "tests", "fortran")
CALLER = os.path.join(CALLER_DIR, "_build", "rtm_caller")
COND_CALLER = os.path.join(CALLER_DIR, "_build", "cond_caller")
SU_CALLER = os.path.join(CALLER_DIR, "_build", "su_caller")
TT_CALLER = os.path.join(CALLER_DIR, "_build", "tt_caller")


def build_fortran_caller(verbose: bool = False):
    """Compiles the Fortran shim against the test fixture module mod_parfile and links
    the test callers (tests/fortran/rtm_caller.f90: the solver calls of one shot;
    tests/fortran/cond_caller.f90: the gather conditioning; tests/fortran/su_caller.f90: the
    SU shot reader; tests/fortran/tt_caller.f90: the Kirchhoff travel times, with the window
    fixture dynamic_window_test.f90) with libpsdm_rtm.so.
    Returns the solver caller's path, or None where no Fortran compiler is installed."""
    if not os.path.exists(FLANG):
        return None
    bdir = os.path.join(CALLER_DIR, "_build")
    os.makedirs(bdir, exist_ok=True)
    rpath = "$ORIGIN/" + os.path.relpath(PKG, bdir)
    common = [os.path.join(CALLER_DIR, "mod_parfile_test.f90"), FORTRAN_SHIM]
    for main, exe in (("rtm_caller.f90", CALLER), ("cond_caller.f90", COND_CALLER), ("su_caller.f90", SU_CALLER),
                      ("tt_caller.f90", TT_CALLER)):
        extra = [os.path.join(CALLER_DIR, "dynamic_window_test.f90")]  # the shim's travel-time entries call it
        cmd = [FLANG, "-O2", "-J", bdir, *common, *extra, os.path.join(CALLER_DIR, main), "-L", PKG, "-lpsdm_rtm",
               f"-Wl,-rpath,{rpath}", "-o", exe]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, cwd=bdir)
    return CALLER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran_caller(verbose="-v" in sys.argv))
