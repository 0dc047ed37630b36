"""ctypes binding of libpsdm_rtm.so (the C ABI in include/psdm_rtm.h).

The library is built in-tree by psdm_amd.build.  There is no fallback: if the shared
library is missing or the HIP runtime has no device, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libpsdm_rtm.so")

_P, _I, _F, _LL = C.c_void_p, C.c_int, C.c_float, C.c_longlong

# exported symbols and their argument types
_HOST_RAM = [_P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _F, _F, _I, _I, _P, _I, _I, _I, _P, _I, _P]
_HOST_DISK = [_P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _F, _F, _I, _I, _P, _I, _I, _I, _I, _P]
SIGNATURES = {
    "psdm_rtm_create": ([C.POINTER(_P), _I], _I),
    "psdm_rtm_destroy": ([_P], _I),
    "psdm_rtm_last_error": ([_P], C.c_char_p),
    "psdm_rtm_forward_ram": (_HOST_RAM, _I),
    "psdm_rtm_backward_ram": (_HOST_RAM, _I),
    "psdm_rtm_forward_disk": (_HOST_DISK, _I),
    "psdm_rtm_backward_disk": (_HOST_DISK, _I),
    "psdm_rtm_forward_device": ([_P, _P, _P, _P, _P, _LL], _I),
    "psdm_rtm_backward_device": ([_P, _P, _P, _P, _P, _LL], _I),
    "psdm_rtm_sync": ([_P], _I),
    "psdm_rtm_stack_image": ([_P, _P, _I, _I, _I, _P, _P, _I], _I),
    "psdm_rtm_normalize": ([_P, _P, _P, _I, _I, _P], _I),
    "psdm_rtm_device_bytes": ([_P], _LL),
    "psdm_rtm_set_launch": ([_P, _I, _I, _I], _I),
    "psdm_rtm_set_stream_width": ([_P, _I], _I),
    "psdm_rtm_set_timing": ([_P, _I], _I),
    "psdm_rtm_set_weights": ([_P, _I], _I),
    "psdm_rtm_step_stats": ([_P, C.POINTER(_LL), C.POINTER(C.c_double)], _I),
    "psdm_rtm_engine_create": ([_P, _P, C.POINTER(_P)], _I),
    "psdm_rtm_engine_destroy": ([_P], _I),
    "psdm_rtm_engine_submit": ([_P, _P, C.POINTER(_LL)], _I),
    "psdm_rtm_engine_run": ([_P, _I, C.POINTER(_I), C.POINTER(_I)], _I),
    "psdm_rtm_engine_poll": ([_P, C.POINTER(_LL), C.POINTER(C.c_double), _I, C.POINTER(_I)], _I),
    "psdm_rtm_engine_drain": ([_P], _I),
    "psdm_rtm_engine_image": ([_P, C.POINTER(_P), C.POINTER(_P)], _I),
    "psdm_rtm_engine_timing": ([_P, C.POINTER(C.c_double), C.POINTER(_LL)], _I),
    "psdm_rtm_condition": ([_P, _P, _I, _P, _P], _I),
    "psdm_rtm_condition_host": ([_P, _P, _I, _P, _P], _I),
    "psdm_rtm_psdm2twt": ([_P, _I, _I, _F, _F, _P, _P, _I, _P, _P, C.POINTER(_I)], _I),
    "psdm_rtm_psdm2twt_noalloc": ([_P, _I, _I, _F, _F, _P, _P, _I, _P, _P, C.POINTER(_I)], _I),
    "psdm_rtm_psdm2twt_timing": ([_P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(_LL)], _I),
    "psdm_rtm_condition_timing": ([_P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(_LL)], _I),
    "psdm_rtm_engine_stats": ([_P, C.POINTER(_LL), C.POINTER(_LL), C.POINTER(_LL), C.POINTER(_LL)], _I),
    "psdm_rtm_engine_memory": ([_P, C.POINTER(_LL), C.POINTER(_LL)], _I),
    "psdm_rtm_taper_snapshots": ([_P, _P, _I, _I], _I),
    "psdm_rtm_su_decode": ([_P, _P, _LL, _I, _I, _I, _I, _I, _P], _I),
    "psdm_rtm_su_header_i32": ([_P, _P, _LL, _I, _I, _I, _I, _P], _I),
    "psdm_rtm_su_read_host": ([_P, C.c_char_p, _LL, _I, _I, _I, _I, _I, _P], _I),
    "psdm_rtm_su_encode": ([_P, _P, _I, _I, _F, _F, _I, _I, _P], _I),
    "psdm_rtm_su_timing": ([_P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(_LL), C.POINTER(_LL)], _I),
    "psdm_rtm_eikonal": ([_P, _P, _I, _I, _F, _F, _I, _P, _P, _P], _I),
    "psdm_rtm_eikonal_host": ([_P, _P, _I, _I, _F, _F, _I, _P, _P, _P], _I),
    "psdm_rtm_set_eikonal_workspace": ([_P, C.c_size_t], _I),
    "psdm_rtm_eikonal_batch": ([_P, _I, _I, _I, C.POINTER(_I)], _I),
    "psdm_rtm_eikonal_division": ([_P, _F, _F, C.POINTER(_I)], _I),
    "psdm_rtm_kirchhoff_image": ([_P, _P, _P, _P, _I, _I, _I, _I, _I, _F, _F, _F, _F, _F, _F, _P], _I),
    "psdm_rtm_kirchhoff_timing": ([_P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(_LL),
                                   C.POINTER(_LL)], _I),
    "psdm_rtm_mcs_geometry": ([_P, _I, _F, _I, _F, _I, _F, _F, _F, _F, _F, _P, _P, _P, C.POINTER(_I), C.POINTER(_I)],
                              _I),
    "psdm_rtm_nav_read": ([C.c_char_p, _I, _I, _F, C.POINTER(_I), C.POINTER(_I), _I, _P, _P, _P, C.POINTER(_I),
                           C.POINTER(_I), C.POINTER(_I)], _I),
}


class Geom(C.Structure):
    """psdm_rtm_geom (include/psdm_rtm.h)."""

    _fields_ = [
        ("ny", _I), ("nx", _I), ("nt", _I), ("dPML", _I), ("fs", _I), ("variant", _I),
        ("store_snap", _I), ("nt_snap", _I), ("dmodel", _F), ("dt", _F), ("NumSou", _I), ("NumRec", _I),
        ("nxSou", _P), ("nySou", _P), ("nxRec", _P), ("nyRec", _P),
    ]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m psdm_amd.build` (there is no CPU fallback)")
        # torch ships its own libamdhip64.so.7 (same soname as /opt/rocm's): whichever
        # loads first serves the whole process, and torch cannot initialise on the other
        # one.  Load torch's first so that torch tensors and this library share it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


class RtmError(RuntimeError):
    pass


class Context:
    """One psdm_rtm_ctx: device buffers, device-resident snapshots and a HIP stream."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = _P()
        rc = self._L.psdm_rtm_create(C.byref(h), device)
        if rc != 0:
            raise RtmError(f"psdm_rtm_create(device={device}) failed with {rc}: no usable HIP device")
        self.h = h
        self.device = device

    def check(self, rc: int, what: str):
        if rc != 0:
            msg = self._L.psdm_rtm_last_error(self.h)
            raise RtmError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self._L.psdm_rtm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # launch controls ------------------------------------------------------------
    def set_launch(self, tile_x: int = 8, use_graph: bool = False, graph_steps: int = 64):
        self.check(self._L.psdm_rtm_set_launch(self.h, tile_x, int(use_graph), graph_steps), "set_launch")

    def set_stream_width(self, stream_w: int):
        self.check(self._L.psdm_rtm_set_stream_width(self.h, stream_w), "set_stream_width")

    def set_weights(self, weights: str):
        """Second-derivative weights of the device solver: "reference" (the default,
        bit-exact with the reference) or "zero_sum" (psdm_rtm_set_weights)."""
        code = {"reference": 0, "zero_sum": 1}[weights]
        self.check(self._L.psdm_rtm_set_weights(self.h, code), "set_weights")

    def set_timing(self, every: int):
        self.check(self._L.psdm_rtm_set_timing(self.h, every), "set_timing")

    def step_stats(self):
        n, ms = _LL(), C.c_double()
        self.check(self._L.psdm_rtm_step_stats(self.h, C.byref(n), C.byref(ms)), "step_stats")
        return n.value, ms.value

    def sync(self):
        self.check(self._L.psdm_rtm_sync(self.h), "sync")

    def device_bytes(self) -> int:
        return int(self._L.psdm_rtm_device_bytes(self.h))
