"""ctypes bindings of the gfx950 libraries and device-memory plumbing.

The product path is the HIP code in bgk3d_gpu_amd/lib/*.so.  If a library is
missing this module raises: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIBDIR = Path(__file__).resolve().parent / "lib"

_libs: dict[str, ctypes.CDLL] = {}


# init-mode flag of bgk_init_host / bgk_init_gpu: unshifted storage (the
# reference's -DNOSHIFT builds), BGK_INIT_UNSHIFTED in include/bgk3d_gpu.h
INIT_UNSHIFTED = 8


class BGKError(RuntimeError):
    pass


def _load(name: str) -> ctypes.CDLL:
    if name not in _libs:
        path = LIBDIR / name
        if not path.exists():
            raise BGKError(f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
        _libs[name] = ctypes.CDLL(str(path))
    return _libs[name]


def real_ctype(prec: str):
    return ctypes.c_float if prec == "sp" else ctypes.c_double


def np_dtype(prec: str):
    return np.float32 if prec == "sp" else np.float64


def v1(prec: str = "sp") -> ctypes.CDLL:
    """The reference-compatible library for `prec` ('sp' or 'dp')."""
    lib = _load("libbgk3d_gpu_sp.so" if prec == "sp" else "libbgk3d_gpu_dp.so")
    if not getattr(lib, "_bgk_typed", False):
        R = real_ctype(prec)
        vp, ci = ctypes.c_void_p, ctypes.c_int
        lib.col_mc_gpu.argtypes = [vp] * 37 + [ci] * 5 + [R] * 12
        lib.col_mc_gpu.restype = None
        lib.col_mc_force_gpu.argtypes = [vp, vp] + [ci] * 5 + [R, R, ci]
        lib.col_mc_les_gpu.argtypes = [vp, vp] + [ci] * 5 + [R, R, R, ci]
        lib.col_mc_force_gpu.restype = None
        for f in ("bcond_periodic_gpu", "bcond_channel_gpu"):
            getattr(lib, f).argtypes = [vp] + [ci] * 5
            getattr(lib, f).restype = None
        lib.bcond_driven_gpu.argtypes = [vp] + [ci] * 5 + [R]
        lib.bcond_driven_gpu.restype = None
        lib.bcond_inflow_gpu.argtypes = [vp] + [ci] * 5 + [R] * 7
        lib.bcond_inflow_gpu.restype = None
        lib.bcond_obs_gpu.argtypes = [vp, vp] + [ci] * 13
        lib.bcond_obs_gpu.restype = None
        lib.gpu_device_sync.argtypes = []
        lib.gpu_device_sync.restype = None
        lib.diagno_gpu.argtypes = [vp] + [ci] * 5 + [R, vp]
        lib.diagno_gpu_noalloc.argtypes = [vp] + [ci] * 5 + [R, vp]
        lib.varm_gpu.argtypes = [vp] + [ci] * 5 + [R, vp, vp, vp, vp]
        lib.draglift_gpu.argtypes = [vp, vp] + [ci] * 9 + [R, vp]
        lib.prof_k_gpu.argtypes = [vp] + [ci] * 7 + [R, vp]
        lib.prof_i_gpu.argtypes = [vp] + [ci] * 7 + [R, vp]
        lib.prof_j_gpu.argtypes = [vp] + [ci] * 7 + [R, vp]
        lib.probe_gpu.argtypes = [vp] + [ci] * 8 + [R, vp]
        lib.bgk_set_stream.argtypes = [vp]
        lib.bgk_dev_malloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        lib.bgk_dev_free.argtypes = [vp]
        lib.bgk_memcpy_h2d.argtypes = [vp, vp, ctypes.c_size_t]
        lib.bgk_memcpy_d2h.argtypes = [vp, vp, ctypes.c_size_t]
        lib.bgk_device_count.argtypes = [ctypes.POINTER(ci)]
        lib.bgk_event_create.argtypes = [ctypes.POINTER(vp)]
        lib.bgk_event_record.argtypes = [vp]
        lib.bgk_event_elapsed_ms.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_float)]
        lib.bgk_event_destroy.argtypes = [vp]
        lib.bgk_stream_copy.argtypes = [vp, vp, ctypes.c_size_t]
        lib._bgk_typed = True
    return lib


def check(rc: int, what: str):
    if rc != 0:
        raise BGKError(f"{what} failed with code {rc}")


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = v1("sp").bgk_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


class DeviceBuffer:
    """A hipMalloc'd buffer owned by Python (freed on close/GC)."""

    def __init__(self, nbytes: int, lib=None):
        self.lib = lib or v1("sp")
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(self.lib.bgk_dev_malloc(ctypes.byref(p), self.nbytes), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    def upload(self, host: np.ndarray, offset_bytes: int = 0):
        host = np.ascontiguousarray(host)
        assert offset_bytes + host.nbytes <= self.nbytes
        check(self.lib.bgk_memcpy_h2d(self.ptr + offset_bytes, host.ctypes.data, host.nbytes), "H2D copy")

    def download(self, host: np.ndarray, offset_bytes: int = 0):
        assert host.flags.c_contiguous and offset_bytes + host.nbytes <= self.nbytes
        check(self.lib.bgk_memcpy_d2h(host.ctypes.data, self.ptr + offset_bytes, host.nbytes), "D2H copy")
        return host

    def close(self):
        if self.ptr:
            self.lib.bgk_dev_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    """hipEvent recorded on the library's ABI stream."""

    def __init__(self, lib):
        self.lib = lib
        p = ctypes.c_void_p()
        check(lib.bgk_event_create(ctypes.byref(p)), "hipEventCreate")
        self.ptr = p.value

    def record(self):
        check(self.lib.bgk_event_record(self.ptr), "hipEventRecord")

    def elapsed_ms(self, later: "Event") -> float:
        ms = ctypes.c_float()
        check(self.lib.bgk_event_elapsed_ms(self.ptr, later.ptr, ctypes.byref(ms)), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            self.lib.bgk_event_destroy(self.ptr)
        except Exception:
            pass
