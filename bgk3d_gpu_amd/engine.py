"""Python binding of the handle API (include/bgk3d_engine.h).

`Engine(case, storage)` owns a lattice on the GPU laid out for gfx950 with
the boundary conditions folded into the fused step.  Storage is 'f32' or
'f64' (bit-identical to the reference's SINGLE / DOUBLE_P builds), or 'f16' /
'bf16' (shifted populations in 16 bits, FP32 collision).  `mixed=1` ('f16')
and `mixed=2` ('f32') are the reference's MIXED1 and MIXED2 builds.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import runtime as rt
from .cases import Case

STORAGE = {"f32": 0, "f64": 1, "f16": 2, "bf16": 3}
BOUNDARY = {"ldc": 0, "periodic": 1, "channel": 2, "inflow": 3}


class BGKConfig(ctypes.Structure):
    _fields_ = [("l", ctypes.c_int), ("m", ctypes.c_int), ("n", ctypes.c_int), ("boundary", ctypes.c_int),
                ("storage", ctypes.c_int), ("rounding", ctypes.c_int), ("svisc", ctypes.c_double),
                ("u0", ctypes.c_double), ("init", ctypes.c_int), ("obstacle", ctypes.c_int),
                ("obs_ic", ctypes.c_int), ("obs_jc", ctypes.c_int), ("obs_kc", ctypes.c_int),
                ("obs_radius", ctypes.c_int), ("u_inflow", ctypes.c_double), ("force_axis", ctypes.c_int),
                ("device", ctypes.c_int), ("seed", ctypes.c_uint), ("les", ctypes.c_int),
                ("cte_s", ctypes.c_double), ("wall_axis", ctypes.c_int), ("mixed", ctypes.c_int),
                ("layout", ctypes.c_int)]


def lib() -> ctypes.CDLL:
    L = rt._load("libbgk3d_engine.so")
    if not getattr(L, "_bgk_typed", False):
        vp, ci, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
        L.bgk_config_default.argtypes = [ctypes.POINTER(BGKConfig), ci, ci, ci]
        L.bgk_config_default.restype = None
        L.bgk_create.argtypes = [ctypes.POINTER(BGKConfig), ctypes.POINTER(vp)]
        L.bgk_destroy.argtypes = [vp]
        L.bgk_step.argtypes = [vp, ci]
        L.bgk_sync.argtypes = [vp]
        L.bgk_time_steps.argtypes = [vp, ci, dp, dp]
        L.bgk_diagno.argtypes = [vp, vp]
        L.bgk_draglift.argtypes = [vp, vp]
        L.bgk_fields.argtypes = [vp, vp, vp, vp, vp]
        L.bgk_get_population.argtypes = [vp, ci, vp]
        L.bgk_prof_k.argtypes = [vp, ci, ci, vp]
        L.bgk_profile.argtypes = [vp, ci, ci, ci, vp]
        L.bgk_probe.argtypes = [vp, ci, ci, ci, vp]
        L.bgk_save_raw.argtypes = [vp, ctypes.c_char_p, ci]
        L.bgk_restore_raw.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ci)]
        L.bgk_save_vel.argtypes = [vp, ctypes.c_char_p, ci]
        L.bgk_write_vtk.argtypes = [vp, ci, ci, ctypes.c_char_p]
        L.bgk_set_population.argtypes = [vp, ci, vp]
        L.bgk_info.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
        L.bgk_last_error.restype = ctypes.c_char_p
        L._bgk_typed = True
    return L


def _check(rc, what):
    if rc != 0:
        msg = lib().bgk_last_error().decode(errors="replace")
        raise rt.BGKError(f"{what}: {msg} (code {rc})")


def make_config(case: Case, storage: str = "f32", rounding: int = 0, obstacle: int | None = None,
                centre=None, radius=None, u_inflow=None, device=0, seed=0, les=None, cte_s=None,
                mixed: int = 0, layout: int = 0) -> BGKConfig:
    """mixed: 1 the reference's MIXED1 form (storage 'f16', reference-order
    FP32 collision), 2 its MIXED2 form (storage 'f32', FP64 collision).
    layout: 0 choose, 1 rows padded to 128-byte lines, 2 compact rows."""
    c = BGKConfig()
    lib().bgk_config_default(ctypes.byref(c), case.l, case.m, case.n)
    c.boundary = BOUNDARY[case.boundary]
    c.storage = STORAGE[storage]
    c.rounding = rounding
    c.svisc = case.svisc
    c.u0 = case.u0
    c.init = case.init
    c.obstacle = (1 if case.obstacle else 0) if obstacle is None else obstacle
    if centre is not None:
        c.obs_ic, c.obs_jc, c.obs_kc = centre
    if radius is not None:
        c.obs_radius = radius
    c.u_inflow = -1.0 if u_inflow is None else u_inflow
    c.force_axis = case.force_axis
    c.device = device
    c.seed = seed
    c.les = 1 if (case.les if les is None else les) else 0
    c.cte_s = case.cte_s if cte_s is None else cte_s
    c.wall_axis = case.wall_axis
    c.mixed = mixed
    c.layout = layout
    return c


class Engine:
    def __init__(self, case: Case, storage: str = "f32", **kw):
        self.case, self.storage = case, storage
        self.L = lib()
        self.cfg = make_config(case, storage, **kw)
        h = ctypes.c_void_p()
        _check(self.L.bgk_create(ctypes.byref(self.cfg), ctypes.byref(h)), "bgk_create")
        self.h = h.value

    def step(self, n: int = 1):
        _check(self.L.bgk_step(self.h, n), "bgk_step")

    def sync(self):
        _check(self.L.bgk_sync(self.h), "bgk_sync")

    def time_steps(self, n: int):
        tot, kern = ctypes.c_double(), ctypes.c_double()
        _check(self.L.bgk_time_steps(self.h, n, ctypes.byref(tot), ctypes.byref(kern)), "bgk_time_steps")
        return tot.value, kern.value

    def diagno(self) -> np.ndarray:
        out = np.zeros(5)
        _check(self.L.bgk_diagno(self.h, out.ctypes.data), "bgk_diagno")
        return out

    def draglift(self) -> np.ndarray:
        out = np.zeros(3)
        _check(self.L.bgk_draglift(self.h, out.ctypes.data), "bgk_draglift")
        return out

    def fields(self):
        """rho, ux, uy, uz as (n, m, l) float64 arrays."""
        c = self.case
        out = [np.empty((c.n, c.m, c.l)) for _ in range(4)]
        _check(self.L.bgk_fields(self.h, *[o.ctypes.data for o in out]), "bgk_fields")
        return out

    def prof_k(self, ic=None, jc=None) -> np.ndarray:
        """(n, 4) array of u, w (y-velocity), v (z-velocity), rho along k."""
        c = self.case
        out = np.zeros((c.n, 4))
        _check(self.L.bgk_prof_k(self.h, c.l // 2 if ic is None else ic, c.m // 2 if jc is None else jc,
                                 out.ctypes.data), "bgk_prof_k")
        return out

    def profile(self, axis: int, c1: int, c2: int) -> np.ndarray:
        """(len, 4) rows of u, w (y-velocity), v (z-velocity), rho along axis
        0 (i at (j, k) = (c1, c2)), 1 (j at (i, k)) or 2 (k at (i, j))."""
        c = self.case
        out = np.zeros(((c.l, c.m, c.n)[axis], 4))
        _check(self.L.bgk_profile(self.h, axis, c1, c2, out.ctypes.data), "bgk_profile")
        return out

    def probe(self, i: int, j: int, k: int) -> np.ndarray:
        """u, v, w, rho at one cell."""
        out = np.zeros(4)
        _check(self.L.bgk_probe(self.h, i, j, k, out.ctypes.data), "bgk_probe")
        return out

    def save_raw(self, path, itime: int):
        _check(self.L.bgk_save_raw(self.h, str(path).encode(), itime), "bgk_save_raw")

    def restore_raw(self, path) -> int:
        it = ctypes.c_int()
        _check(self.L.bgk_restore_raw(self.h, str(path).encode(), ctypes.byref(it)), "bgk_restore_raw")
        return it.value

    def save_vel(self, path, itime: int):
        _check(self.L.bgk_save_vel(self.h, str(path).encode(), itime), "bgk_save_vel")

    def write_vtk(self, kind: int, index: int, path):
        """kind 0 xy, 1 xz, 2 yz velocity plane; 3 xy vorticity plane."""
        _check(self.L.bgk_write_vtk(self.h, kind, index, str(path).encode()), "bgk_write_vtk")

    def population(self, d: int) -> np.ndarray:
        c = self.case
        out = np.empty((c.n, c.m, c.l))
        _check(self.L.bgk_get_population(self.h, d, out.ctypes.data), "bgk_get_population")
        return out

    def populations(self) -> np.ndarray:
        return np.stack([self.population(d) for d in range(19)])

    def set_population(self, d: int, values: np.ndarray):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.shape == (self.case.n, self.case.m, self.case.l)
        _check(self.L.bgk_set_population(self.h, d, v.ctypes.data), "bgk_set_population")

    def info(self) -> dict:
        out = (ctypes.c_longlong * 6)()
        _check(self.L.bgk_info(self.h, out), "bgk_info")
        return dict(nx=out[0], ny=out[1], nz=out[2], bytes=out[3], esize=out[4], steps=out[5])

    def close(self):
        if self.h:
            self.L.bgk_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
