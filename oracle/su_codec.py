"""SU trace files and shot navigation -- TEST INFRASTRUCTURE ONLY (the oracle).

CPU restatement with numpy of:

* ``decode`` -- get_shot_data (get_mcs_data.f90:150-214) / get_su_data (auxiliar.f90:9-44):
  trace j starts at byte j*(60+ns)*4 (:171-173), its samples are words 60..60+nt-1 (:177-181),
  read as big- or little-endian IEEE FP32 (the unit's CONVERT, :303-312); ``shift`` moves
  sample i to i+shift as the delayed gather does (:184-204; samples it leaves unwritten are
  0 here);
* ``header_i32`` -- the INTEGER*4 at a 1-based byte position of each trace header
  (get_pos_channels :457, :467);
* ``geometry`` -- get_pos_channels (:418-510) and get_geometry_mcs (:128-148) in FP32;
* ``read_nav`` -- read_nav_shot_mcs (:30-126);
* ``encode`` -- the SU file write_psdm's Seismic Unix pipeline produces (auxiliar.f90:
  275-299): a2b of the samples, suaddhead ns=nj (tracl = tracr = 1, 2, ...), sushw d1=dj
  f1=0 d2=di f2=0 trid, suchw cdp = (-di + di*tracl)/1.  Header layout per Seismic Unix's
  segy.h (tracl 0, tracr 4, cdp 20, trid 28, ns 114, d1 180, f1 184, d2 188, f2 192).

Pinned by tests/test_su_oracle.py: decode, header_i32, geometry and read_nav against the
reference's own routines compiled into oracle/_ref (open_shot_mcs + get_shot_data,
get_pos_channels + get_geometry_mcs, read_nav_shot_mcs); encode by reading its files back
through the reference's get_su_data and by the header table written out independently in
the test.  Seismic Unix itself is not in this image, so the header defaults it would add
beyond the fields named above are parity unpinned.
"""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np

from . import pyoracle as _po

F32 = np.float32
HDR_WORDS = 60


def _dt(big: bool, kind: str) -> np.dtype:
    return np.dtype((">" if big else "<") + kind)


def decode(buf, ntr: int, ns: int, nt: int, big: bool = True, shift: int = 0) -> np.ndarray:
    """(ntr, nt) FP32 gather, trace-major."""
    w = np.frombuffer(memoryview(buf), dtype=_dt(big, "f4"), count=ntr * (HDR_WORDS + ns)).reshape(ntr, HDR_WORDS + ns)
    raw = w[:, HDR_WORDS:HDR_WORDS + nt].astype(F32)
    if shift == 0:
        return raw
    out = np.zeros_like(raw)
    if shift > 0:
        if shift < nt:
            out[:, shift:] = raw[:, :nt - shift]
    elif -shift < nt:
        out[:, :nt + shift] = raw[:, -shift:]
    return out


def header_i32(buf, ntr: int, ns: int, pos: int, big: bool = True) -> np.ndarray:
    b = np.frombuffer(memoryview(buf), dtype=np.uint8)
    out = np.empty(ntr, np.int32)
    fmt = ">i" if big else "<i"
    for j in range(ntr):
        o = j * (HDR_WORDS + ns) * 4 + pos - 1
        out[j] = struct.unpack(fmt, b[o:o + 4].tobytes())[0]
    return out


def geometry(offsets, nrec: int, pos_shot: float, offset_header: int, offset_unit: float, dir_boat: int,
             near_offset: float, drec: float, dd: float, shot_depth: float, streamer_depth: float):
    """(pos_rec, nxRec, nyRec, nxSou, nySou), FP32 arithmetic in the reference's order."""
    ps, dd, dirf = F32(pos_shot), F32(dd), F32(dir_boat)
    if offset_header == 1:
        a = np.abs(np.asarray(offsets, np.int64)).astype(F32)
        off = a * F32(offset_unit) if offset_unit > 0 else a / F32(abs(F32(offset_unit)))
    else:
        off = F32(near_offset) + np.arange(nrec).astype(F32) * F32(drec)
    pos_rec = (ps - (dirf * off.astype(F32)).astype(F32)).astype(F32)
    nxr = 1 + np.floor((pos_rec / dd).astype(F32)).astype(np.int64)
    nyr = np.full(nrec, 1 + int(np.floor(F32(streamer_depth) / dd)), np.int64)
    return (pos_rec, nxr.astype(np.int32), nyr.astype(np.int32), 1 + int(np.floor(ps / dd)),
            1 + int(np.floor(F32(shot_depth) / dd)))


def read_nav(path: str, cols: int, dir_boat: int, d_shot_model: float, shot_i: int = 0, shot_f: int = 0):
    """(ids, pos, bat, shot_i, shot_f, line_i, line_f)."""
    rows = [ln.replace(",", " ").split() for ln in open(path).read().splitlines()]
    nv = {11: 1, 111: 2, 1111: 3}[cols]
    ids = np.array([int(r[0]) for r in rows], np.int32)
    v = np.array([[float(x) for x in r[1:1 + nv]] + [0.0] * (3 - nv) for r in rows]).astype(F32)
    xs, ys, bt = v[:, 0], v[:, 1], v[:, 2]
    r0 = 0 if dir_boat == 1 else len(rows) - 1
    dx, dy = (xs - xs[r0]).astype(F32), (ys - ys[r0]).astype(F32)
    pos = (F32(d_shot_model) + np.sqrt((dx * dx).astype(F32) + (dy * dy).astype(F32)).astype(F32)).astype(F32)
    line_i = line_f = 0
    for r, sid in enumerate(ids):
        if r == 0 and shot_i == 0:
            shot_i = int(sid)
        if r == len(ids) - 1 and shot_f == 0:
            shot_f = int(sid)
        if shot_i > 0 and shot_i == sid:
            line_i = r + 1
        if shot_f > 0 and shot_f == sid:
            line_f = r + 1
    return ids, pos, np.abs(bt), shot_i, shot_f, line_i, line_f


def su_parsed(x: float) -> float:
    """The value Seismic Unix parses from the reference's list-directed write of a REAL(4)."""
    return float(f"{float(F32(x)):.8E}")


def encode(img, di: float, dj: float, trid: int, big: bool = True) -> bytes:
    """SU bytes of img (ni, nj) trace-major."""
    img = np.ascontiguousarray(img, F32)
    ni, nj = img.shape
    e = ">" if big else "<"
    dis = su_parsed(di)
    out = bytearray()
    for j in range(ni):
        h = bytearray(240)
        tracl = j + 1
        struct.pack_into(e + "i", h, 0, tracl)
        struct.pack_into(e + "i", h, 4, tracl)
        struct.pack_into(e + "i", h, 20, int((-dis + dis * tracl + 0.0 * tracl) / 1.0))
        struct.pack_into(e + "h", h, 28, trid)
        struct.pack_into(e + "H", h, 114, nj)
        struct.pack_into(e + "ffff", h, 180, float(F32(dj)), 0.0, float(F32(di)), 0.0)
        out += h
        out += img[j].astype(_dt(big, "f4")).tobytes()
    return bytes(out)


# --- the reference's routines (oracle/_ref) -------------------------------------------

def _bind(name, argtypes):
    L = _po.ref_lib()
    fn = getattr(L, name)
    if not getattr(fn, "_bound", False):
        fn.argtypes = argtypes
        fn.restype = None
        fn._bound = True
    return fn


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def ref_read_shot(path: str, big: bool, nrec: int, ntmax: int, nt: int) -> np.ndarray:
    """open_shot_mcs + get_shot_data: (nrec, nt) trace-major."""
    fn = _bind("ref_su_read_shot", [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p])
    raw = np.zeros((nt, nrec), F32, order="F")
    fn(path.encode(), int(big), nrec, ntmax, nt, _p(raw))
    return np.ascontiguousarray(raw.T)


def ref_geometry(path: str, big: bool, nrec: int, ntmax: int, byte_shot: int, offset_header: int,
                 offset_unit: float, dir_boat: int, near_offset: float, drec: float, pos_shot: float, dd: float,
                 shot_depth: float, streamer_depth: float):
    """get_pos_channels + get_geometry_mcs: (shot_id, offsets_su, pos_rec, nxRec, nyRec, nxSou, nySou)."""
    fn = _bind("ref_su_geometry", [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                   C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p])
    sid, nxs, nys = (np.zeros(1, np.int32) for _ in range(3))
    offs, posr = np.zeros(nrec, F32), np.zeros(nrec, F32)
    nxr, nyr = np.zeros(nrec, np.int32), np.zeros(nrec, np.int32)
    fn(path.encode(), int(big), nrec, ntmax, byte_shot, offset_header, offset_unit, dir_boat, near_offset, drec,
       pos_shot, dd, shot_depth, streamer_depth, _p(sid), _p(offs), _p(posr), _p(nxs), _p(nys), _p(nxr), _p(nyr))
    return int(sid[0]), offs, posr, nxr, nyr, int(nxs[0]), int(nys[0])


def ref_read_nav(path: str, cols: int, dir_boat: int, d_shot_model: float, n: int):
    """read_nav_shot_mcs: (ids, pos, bat, line_i, line_f)."""
    fn = _bind("ref_read_nav", [C.c_char_p, C.c_int, C.c_int, C.c_float, C.c_int, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_void_p, C.c_void_p])
    ids, pos, bat = np.zeros(n, np.int32), np.zeros(n, F32), np.zeros(n, F32)
    li, lf = np.zeros(1, np.int32), np.zeros(1, np.int32)
    fn(path.encode(), cols, dir_boat, d_shot_model, n, _p(ids), _p(pos), _p(bat), _p(li), _p(lf))
    return ids, pos, bat, int(li[0]), int(lf[0])


def ref_get_su_data(path: str, big: bool, ni: int, ns: int, ns_max: int) -> np.ndarray:
    """get_su_data: (ni, ns_max) trace-major."""
    fn = _bind("ref_get_su_data", [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p])
    out = np.zeros((ns_max, ni), F32, order="F")
    fn(path.encode(), int(big), ni, ns, ns_max, _p(out))
    return np.ascontiguousarray(out.T)
