"""ctypes front end of the oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (the C restatement of the reference solver, FP32 and
FP64) and, where it was built, ``oracle/_ref/libpsdm_ref.so`` (the reference Fortran
solvers compiled from their own sources).  Only tests/, ``__graft_entry__.smoke()``
and bench.py's ``cpu_baseline`` leg import this module, and only as the checker.

Array conventions follow the reference: Fortran (column-major) arrays, shapes
``S(nt, NumSou)``, ``Sadj(nt, NumRec)``, ``ctmp(ny, nx)``, ``Fields(ny, nx, nt_snap)``,
``Rec_point(nt, NumRec)``, ``kernel(ny, nx)``; geometry indices are 1-based
(solver_ac_forward_ram.f90:4-6, solver_ac_backward_ram.f90:4-6).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libpsdm_ref.so")

_P = C.c_void_p
_I = C.c_int
_F = C.c_float

_oracle = None
_ref = None


def build(ref: bool = False) -> None:
    """Compile the oracle (and, if asked and the reference is mounted, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        sig = [_I] * 7 + [_P] * 5 + [_F, _F, _I, _I, _P, _I, _I, _I, _P, _P, _P]
        for name in ("rtm_oracle_solver_f32", "rtm_oracle_solver_f64"):
            f = getattr(L, name)
            f.argtypes = sig
            f.restype = _I
        for name in ("rtm_oracle_pml_tables_f32", "rtm_oracle_pml_tables_f64"):
            f = getattr(L, name)
            f.argtypes = [_I, _I, _F, _F, _P, _P, _P, _P]
            f.restype = None
        L.rtm_oracle_fd_coef_f64.argtypes = [_I, _P]
        L.rtm_oracle_fd_coef_f32.argtypes = [_I, _P]
        L.rtm_oracle_deriv_f64.argtypes = [_I, _I, _I, _P, _P]
        L.rtm_oracle_deriv_f32.argtypes = [_I, _I, _I, _P, _P]
        _oracle = L
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        # modules.f90 references ascii_art (auxiliar.f90) from parfile error paths that
        # the solvers never reach: bind lazily instead of linking more of the driver.
        L = C.CDLL(REF_SO, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        ram = [_I] * 4 + [_P] * 5 + [_F, _F, _I, _I, _P, _I, _I, _I, _P, _P]
        L.ref_forward_ram.argtypes = ram
        L.ref_backward_ram.argtypes = ram
        disk = [_I] * 4 + [_P] * 5 + [_F, _F, _I, _I, _P, _I, C.c_char_p, _I, _I, _I, _P]
        L.ref_forward_disk.argtypes = disk
        L.ref_backward_disk.argtypes = disk
        L.ref_pml_coeffs.argtypes = [_I, _F, _F, _P, _P, _P, _P]
        L.ref_fd_coef.argtypes = [_I, _P]
        L.ref_deriv.argtypes = [_I, _I, _I, _P, _P]
        for f in (L.ref_forward_ram, L.ref_backward_ram, L.ref_forward_disk, L.ref_backward_disk, L.ref_pml_coeffs,
                  L.ref_fd_coef, L.ref_deriv):
            f.restype = None
        _ref = L
    return _ref


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).ravel())


def _f32F(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float32))


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def pml_tables(dPML: int, dt: float, d: float, mode: int = 0, precision: str = "f64"):
    """(fA, fB, cornersA, cornersB); corners stacked (4, dn, dn) as ld, rd, lu, ru, each
    indexed [row j, col i] (solver_PML_coeffs.f90:1-75)."""
    dn = dPML + 6
    dt_ = np.float64 if precision == "f64" else np.float32
    fA = np.zeros(dn, dt_)
    fB = np.zeros(dn, dt_)
    cA = np.zeros((4, dn, dn), dt_)  # memory: 4 Fortran (dn,dn) arrays back to back
    cB = np.zeros((4, dn, dn), dt_)
    fn = lib().rtm_oracle_pml_tables_f64 if precision == "f64" else lib().rtm_oracle_pml_tables_f32
    fn(mode, dPML, dt, d, _ptr(fA), _ptr(fB), _ptr(cA), _ptr(cB))
    # each (dn,dn) block was written column-major: transpose to [row, col]
    return fA, fB, cA.transpose(0, 2, 1).copy(), cB.transpose(0, 2, 1).copy()


def ref_pml_tables(dPML: int, dt: float, d: float):
    """PML_coeffs of the reference itself (oracle/_ref); same layout as pml_tables."""
    dn = dPML + 6
    fA = np.zeros(dn, np.float32)
    fB = np.zeros(dn, np.float32)
    cA = np.zeros((4, dn, dn), np.float32)
    cB = np.zeros((4, dn, dn), np.float32)
    ref_lib().ref_pml_coeffs(dPML, dt, d, _ptr(fA), _ptr(fB), _ptr(cA), _ptr(cB))
    return fA, fB, cA.transpose(0, 2, 1).copy(), cB.transpose(0, 2, 1).copy()


def ref_fd_coef(num: int):
    c = np.zeros(4, np.float32)
    ref_lib().ref_fd_coef(num, _ptr(c))
    return c


def ref_deriv(which: int, p: np.ndarray) -> np.ndarray:
    """A derivative pass of the reference itself (FP32), p indexed [row, col]."""
    pf = np.asfortranarray(p, dtype=np.float32)
    d = np.zeros_like(pf, order="F")
    ref_lib().ref_deriv(which, pf.shape[0], pf.shape[1], _ptr(pf), _ptr(d))
    return d


def ref_ricker_mcs(nt: int, dt: float, f0: float, t0: float) -> np.ndarray:
    """The reference's Ricker_mcs (get_mcs_data.f90:318-351) with nt_prop = nt,
    dt_prop = dt, freq_ricker = f0, delay_ricker = t0."""
    L = ref_lib()
    if not getattr(L, "_ricker_bound", False):
        L.ref_ricker_mcs.argtypes = [_I, _F, _F, _F, _P]
        L.ref_ricker_mcs.restype = None
        L._ricker_bound = True
    s = np.zeros(nt, np.float32)
    L.ref_ricker_mcs(nt, dt, f0, t0, _ptr(s))
    return s


def fd_coef(num: int, precision: str = "f64"):
    dt_ = np.float64 if precision == "f64" else np.float32
    c = np.zeros(4, dt_)
    (lib().rtm_oracle_fd_coef_f64 if precision == "f64" else lib().rtm_oracle_fd_coef_f32)(num, _ptr(c))
    return c


def deriv(which: int, p: np.ndarray, precision: str = "f64") -> np.ndarray:
    """One full-grid derivative pass (solver_space_derivs_clean.f90): which 0 = d2/dx2,
    1 = d2/dy2, 2 = d/dx, 3 = d/dy, 4 / 5 = the zero-sum forms of 0 / 1 (the opt-in
    zero-sum weights mode); p indexed [row, col]."""
    dt_ = np.float64 if precision == "f64" else np.float32
    pf = np.asfortranarray(p, dtype=dt_)
    d = np.zeros_like(pf, order="F")
    fn = lib().rtm_oracle_deriv_f64 if precision == "f64" else lib().rtm_oracle_deriv_f32
    fn(which, pf.shape[0], pf.shape[1], _ptr(pf), _ptr(d))
    return d


def _call(impl, iprop, variant, dPML, fs, nxSou, nySou, nxRec, nyRec, src, dmodel, dt, ctmp, store_snap, nt_snap,
          nt, Fields=None, kernel=None, table_mode=0, precision="f32", shotid=1, folder=None):
    nxSou, nySou, nxRec, nyRec = _i32(nxSou), _i32(nySou), _i32(nxRec), _i32(nyRec)
    NumSou, NumRec = nxSou.size, nxRec.size
    ctmp = _f32F(ctmp)
    ny, nx = ctmp.shape
    src = _f32F(src).reshape((nt, -1), order="F")
    rdt = np.float64 if precision == "f64" else np.float32
    Rec = None
    if iprop == 1:
        Rec = np.zeros((nt, NumRec), rdt, order="F")
        if Fields is None and (variant == 0 or impl == "oracle"):
            Fields = np.zeros((ny, nx, nt_snap), rdt, order="F")
    else:
        kernel = np.zeros((ny, nx), rdt, order="F") if kernel is None else np.asfortranarray(kernel, dtype=rdt)
    if impl == "oracle":
        fn = lib().rtm_oracle_solver_f64 if precision == "f64" else lib().rtm_oracle_solver_f32
        if Fields is not None:
            Fields = np.asfortranarray(Fields, dtype=rdt)
        rc = fn(iprop, variant, table_mode, dPML, fs, NumSou, NumRec, _ptr(nxSou), _ptr(nySou), _ptr(nxRec),
                _ptr(nyRec), _ptr(src), dmodel, dt, ny, nx, _ptr(ctmp), store_snap, nt_snap, nt, _ptr(Fields),
                _ptr(Rec), _ptr(kernel))
        if rc == -2:
            raise ValueError("grid outside the solver's defined range (the _ram variant needs nx >= dPML+6)")
        if rc != 0:
            raise RuntimeError("oracle solver failed (allocation)")
    else:
        L = ref_lib()
        args = (dPML, fs, NumSou, NumRec, _ptr(nxSou), _ptr(nySou), _ptr(nxRec), _ptr(nyRec), _ptr(src), dmodel, dt,
                ny, nx, _ptr(ctmp))
        if variant == 0:
            Fields = np.asfortranarray(Fields, dtype=np.float32)
            if iprop == 1:
                L.ref_forward_ram(*args, store_snap, nt_snap, nt, _ptr(Fields), _ptr(Rec))
            else:
                L.ref_backward_ram(*args, store_snap, nt_snap, nt, _ptr(Fields), _ptr(kernel))
        else:
            fold = (folder.rstrip("/") + "/").encode()
            if iprop == 1:
                L.ref_forward_disk(*args, shotid, fold, store_snap, nt_snap, nt, _ptr(Rec))
            else:
                L.ref_backward_disk(*args, shotid, fold, store_snap, nt_snap, nt, _ptr(kernel))
    return (Rec, Fields) if iprop == 1 else kernel


def forward(nxSou, nySou, nxRec, nyRec, S, dmodel, dt, ctmp, store_snap, nt_snap, nt, dPML, fs, variant=0,
            table_mode=0, precision="f32", impl="oracle", **kw):
    """solver_ac_forward_{ram,disk}: returns (Rec_point(nt,NumRec), Fields(ny,nx,nt_snap) or None)."""
    return _call(impl, 1, variant, dPML, fs, nxSou, nySou, nxRec, nyRec, S, dmodel, dt, ctmp, store_snap, nt_snap,
                 nt, table_mode=table_mode, precision=precision, **kw)


def backward(nxSou, nySou, nxRec, nyRec, Sadj, dmodel, dt, ctmp, store_snap, nt_snap, nt, dPML, fs, Fields=None,
             kernel=None, variant=0, table_mode=0, precision="f32", impl="oracle", **kw):
    """solver_ac_backward_{ram,disk}: returns kernel(ny,nx) (accumulated onto ``kernel``)."""
    return _call(impl, 2, variant, dPML, fs, nxSou, nySou, nxRec, nyRec, Sadj, dmodel, dt, ctmp, store_snap,
                 nt_snap, nt, Fields=Fields, kernel=kernel, table_mode=table_mode, precision=precision, **kw)
