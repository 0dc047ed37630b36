"""SU trace files and shot navigation without Seismic Unix (psdm_rtm_su_*, psdm_rtm_mcs_geometry,
psdm_rtm_nav_read).

The reader replaces get_shot_data (get_mcs_data.f90:150-214) and get_su_data
(auxiliar.f90:9-44): the file's bytes go to the device as they are and one kernel turns
them into the trace-major FP32 gather that psdm_amd.conditioning takes.  The writer replaces
the Seismic Unix pipeline of write_psdm (auxiliar.f90:275-299).  Geometry and navigation
(get_pos_channels, get_geometry_mcs, read_nav_shot_mcs) are host functions of the library.
Device arrays are torch tensors; an image ``(nz, nx)`` in Fortran order is a contiguous
``(nx, nz)`` tensor, one trace per row.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import Context, lib

HDR_BYTES = 240
BYTE_FLDR, BYTE_OFFSET = 9, 37  # modules.f90:20,23: shot number and offset header words


def load_bytes(path: str, device="cuda:0", offset: int = 0, count: int = -1):
    """The file's bytes (from `offset`, `count` of them or all) as a uint8 device tensor,
    staged through pinned host memory."""
    import torch

    a = np.fromfile(path, dtype=np.uint8, count=count, offset=offset)
    h = torch.empty(a.size, dtype=torch.uint8, pin_memory=True)
    h.numpy()[:] = a
    return h.to(device)


def _check_bytes(t):
    import torch

    if t.dtype != torch.uint8 or not t.is_contiguous() or t.dim() != 1:
        raise ValueError("SU bytes must be a contiguous 1-D uint8 tensor")
    _ready(t)


def _ready(t):
    """The library runs on its own stream: finish torch's pending work on t's device first."""
    import torch

    torch.cuda.current_stream(t.device).synchronize()


def decode(ctx: Context, d_su, ntr: int, ns: int, nt: int | None = None, big_endian: bool = True, shift: int = 0,
           out=None):
    """(ntr, nt) float32 gather from SU bytes of ntr traces of ns samples."""
    import torch

    _check_bytes(d_su)
    nt = ns if nt is None else nt
    if out is None:
        out = torch.empty((ntr, nt), dtype=torch.float32, device=d_su.device)
    ctx.check(ctx._L.psdm_rtm_su_decode(ctx.h, C.c_void_p(d_su.data_ptr()), d_su.numel(), ntr, ns, nt,
                                        int(big_endian), shift, C.c_void_p(out.data_ptr())), "psdm_rtm_su_decode")
    return out


def header_i32(ctx: Context, d_su, ntr: int, ns: int, pos: int, big_endian: bool = True):
    """int32 tensor: the header word at 1-based byte `pos` of each trace."""
    import torch

    _check_bytes(d_su)
    out = torch.empty(ntr, dtype=torch.int32, device=d_su.device)
    ctx.check(ctx._L.psdm_rtm_su_header_i32(ctx.h, C.c_void_p(d_su.data_ptr()), d_su.numel(), ntr, ns,
                                            int(big_endian), pos, C.c_void_p(out.data_ptr())),
              "psdm_rtm_su_header_i32")
    return out


def delay_shift(delay_time_data: float, dt_mcs: float) -> int:
    """id = delay_time_data/dt_mcs (get_mcs_data.f90:185), FP32 quotient truncated."""
    return int(np.float32(delay_time_data) / np.float32(dt_mcs))


def encode(ctx: Context, d_img, di: float, dj: float, trid: int = 1, big_endian: bool = True, out=None):
    """SU bytes (uint8 device tensor) of d_img (ni, nj) float32: trace i = row i."""
    import torch

    if d_img.dtype != torch.float32 or not d_img.is_contiguous() or d_img.dim() != 2:
        raise ValueError("d_img must be a contiguous (ni, nj) float32 tensor")
    _ready(d_img)
    ni, nj = d_img.shape
    n = ni * (HDR_BYTES + 4 * nj)
    if out is None:
        out = torch.empty(n, dtype=torch.uint8, device=d_img.device)
    elif out.numel() < n:
        raise ValueError(f"out holds {out.numel()} bytes, {n} needed")
    ctx.check(ctx._L.psdm_rtm_su_encode(ctx.h, C.c_void_p(d_img.data_ptr()), ni, nj, di, dj, trid, int(big_endian),
                                        C.c_void_p(out.data_ptr())), "psdm_rtm_su_encode")
    return out[:n]


def write_su(path: str, d_su) -> None:
    d_su.cpu().numpy().tofile(path)


def timing(ctx: Context):
    """(decode ms, encode ms, decodes, encodes) per call since ctx.set_timing(n != 0)."""
    a, b, na, nb = C.c_double(), C.c_double(), C.c_longlong(), C.c_longlong()
    ctx.check(ctx._L.psdm_rtm_su_timing(ctx.h, C.byref(a), C.byref(b), C.byref(na), C.byref(nb)),
              "psdm_rtm_su_timing")
    return a.value, b.value, na.value, nb.value


def mcs_geometry(offsets, nrec: int, pos_shot: float, offset_header: int = 0, offset_unit: float = 1.0,
                 dir_boat: int = 1, near_offset: float = 0.0, drec: float = 0.0, dd: float = 10.0,
                 shot_depth: float = 0.0, streamer_depth: float = 0.0):
    """(pos_rec, nxRec, nyRec, nxSou, nySou) of one shot (host)."""
    L = lib()
    offs = None if offsets is None else np.ascontiguousarray(offsets, np.int32)
    if offset_header == 1 and (offs is None or offs.size < nrec):
        raise ValueError("offset_header = 1 needs nrec header offsets")
    pos = np.zeros(nrec, np.float32)
    nxr, nyr = np.zeros(nrec, np.int32), np.zeros(nrec, np.int32)
    nxs, nys = C.c_int(), C.c_int()
    rc = L.psdm_rtm_mcs_geometry(None if offs is None else offs.ctypes.data_as(C.c_void_p), nrec, pos_shot,
                                 offset_header, offset_unit, dir_boat, near_offset, drec, dd, shot_depth,
                                 streamer_depth, pos.ctypes.data_as(C.c_void_p), nxr.ctypes.data_as(C.c_void_p),
                                 nyr.ctypes.data_as(C.c_void_p), C.byref(nxs), C.byref(nys))
    if rc != 0:
        raise ValueError("psdm_rtm_mcs_geometry: invalid arguments (offset_header 0/1, offset_unit != 0, dd > 0)")
    return pos, nxr, nyr, nxs.value, nys.value


def read_nav(path: str, cols: int = 11, dir_boat: int = 1, d_shot_model: float = 0.0, shot_i: int = 0,
             shot_f: int = 0):
    """(ids, pos, bat, shot_i, shot_f, line_i, line_f) of a navigation file (host)."""
    L = lib()
    si, sf, n, li, lf = C.c_int(shot_i), C.c_int(shot_f), C.c_int(), C.c_int(), C.c_int()
    args = (path.encode(), cols, dir_boat, d_shot_model)
    if L.psdm_rtm_nav_read(*args, C.byref(si), C.byref(sf), 0, None, None, None, C.byref(n), C.byref(li),
                           C.byref(lf)) != 0:
        raise ValueError(f"psdm_rtm_nav_read: cannot read {path!r} as a {cols}-column navigation file")
    ids, pos, bat = np.zeros(n.value, np.int32), np.zeros(n.value, np.float32), np.zeros(n.value, np.float32)
    si, sf = C.c_int(shot_i), C.c_int(shot_f)
    if L.psdm_rtm_nav_read(*args, C.byref(si), C.byref(sf), n.value, ids.ctypes.data_as(C.c_void_p),
                           pos.ctypes.data_as(C.c_void_p), bat.ctypes.data_as(C.c_void_p), C.byref(n), C.byref(li),
                           C.byref(lf)) != 0:
        raise ValueError(f"psdm_rtm_nav_read: cannot read {path!r}")
    return ids, pos, bat, si.value, sf.value, li.value, lf.value
