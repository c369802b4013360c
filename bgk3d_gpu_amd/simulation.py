"""Host driver over the reference-compatible C-ABI (v1).

This mirrors the reference's program flow (SRC/bgk3d.F90:75-246): setup
allocates two lattices of 19 arrays in the reference layout
(SRC/alloca.F90:150-231, poisoned with huge()), initialize builds the
boundary flags, the obstacle and f = f_eq, hencol sets omega/fgrad, and each
time step calls boundaries -> col_MC -> a<->b pointer swap
(SRC/bgk3d.F90:169-196, SRC/col_MC.F90:539-596).  Every call goes through the
same symbols the reference's Fortran binds, so this is also the executable
specification of how a host drives the library.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import runtime as rt
from .cases import Case, sphere_obstacle

P_CONSTS = {"sp": np.float32, "dp": np.float64}


class Simulation:
    """poison: fill every array with huge() first (the reference's alloca);
    device_init: write f = f_eq with bgk_init_gpu instead of computing it on
    the host and uploading it (same bits; for lattices of many GB)."""

    def __init__(self, case: Case, prec: str = "sp", poison: bool = True, device_init: bool = False):
        if prec not in ("sp", "dp"):
            raise ValueError("prec must be 'sp' or 'dp'")
        self.case, self.prec = case, prec
        self.lib = rt.v1(prec)
        self.R = P_CONSTS[prec]
        self.Rc = rt.real_ctype(prec)
        c = case
        self.nz, self.ny, self.nx = c.shape
        self.count = self.nx * self.ny * self.nz
        nbytes = self.count * np.dtype(self.R).itemsize
        # 19 + 18 arrays: b19 is never used by the fused step (direction 19
        # is updated in place)
        self._a = [rt.DeviceBuffer(nbytes, self.lib) for _ in range(19)]
        self._b = [rt.DeviceBuffer(nbytes, self.lib) for _ in range(18)]
        self.pa = (ctypes.c_void_p * 19)(*[x.ptr for x in self._a])
        self.pb = (ctypes.c_void_p * 19)(*([x.ptr for x in self._b] + [self._a[18].ptr]))
        if poison:
            huge = np.full(self.count, np.finfo(self.R).max, dtype=self.R)
            for buf in self._a + self._b:
                buf.upload(huge)
        self.omega, self.fgrad = c.hencol(prec)
        self.force = c.lid_force(prec)
        if c.noshift and (c.les or c.boundary == "channel"):
            raise ValueError("unshifted storage runs through col_mc_gpu: no channel force or LES form")
        # storage convention: cte1 of the collide and the diagnostics, and of
        # init and the inflow face (SRC/col_MC.F90:139-144, SRC/init.F90:71-75)
        self.cte1 = 0.0 if c.noshift else 1.0
        self.obs_buf = None
        if c.obstacle:
            obs, self.box, self.radius, self.centre = sphere_obstacle(c.l, c.m, c.n)
            self.obs_host = obs
            self.obs_buf = rt.DeviceBuffer(obs.nbytes, self.lib)
            self.obs_buf.upload(obs)
        self.itime = 0
        if device_init:
            self.init_device()
        else:
            self.set_state(self.initial_state())

    # ---- state -----------------------------------------------------------
    def initial_state(self) -> np.ndarray:
        """f = f_eq on every cell incl. halo, computed on the host as the
        reference does (bgk_init_host); padding keeps the huge() poison."""
        c = self.case
        host = np.full((19,) + c.shape, np.finfo(self.R).max, dtype=self.R)
        ptrs = (ctypes.c_void_p * 19)(*[host[d].ctypes.data for d in range(19)])
        self.lib.bgk_init_host.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 6 + [self.Rc]
        mode = c.init | (rt.INIT_UNSHIFTED if c.noshift else 0)
        rt.check(self.lib.bgk_init_host(ptrs, c.l, c.m, c.n, self.nx, self.ny, mode, self.Rc(c.u00)),
                 "bgk_init_host")
        return host

    def init_device(self, mode=None, u00=None):
        """Write the initial state straight into the device arrays with
        bgk_init_gpu (bit-identical to initial_state(); padding untouched)."""
        c = self.case
        mode = (c.init | (rt.INIT_UNSHIFTED if c.noshift else 0)) if mode is None else mode
        u00 = c.u00 if u00 is None else u00
        self.lib.bgk_init_gpu.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 6 + [self.Rc]
        rt.check(self.lib.bgk_init_gpu(self.pa, c.l, c.m, c.n, self.nx, self.ny, mode, self.Rc(u00)), "bgk_init_gpu")

    def set_state(self, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=self.R)
        assert host.shape == (19,) + self.case.shape
        for d in range(19):
            rt.check(self.lib.bgk_memcpy_h2d(self.pa[d], host[d].ctypes.data, host[d].nbytes), "H2D")

    def state(self) -> np.ndarray:
        """Current lattice a01..a19 (halo included) as (19, nz, ny, nx)."""
        self.sync()
        out = np.empty((19,) + self.case.shape, dtype=self.R)
        for d in range(19):
            rt.check(self.lib.bgk_memcpy_d2h(out[d].ctypes.data, self.pa[d], out[d].nbytes), "D2H")
        return out

    # ---- one time step: boundaries -> collision -> swap ------------------
    def boundaries(self):
        c, L, R = self.case, self.lib, self.Rc
        args = (self.pa, c.l, c.m, c.n, self.nx, self.ny)
        if c.boundary == "ldc":
            L.bcond_driven_gpu(*args, R(self.force))
        elif c.boundary == "periodic":
            L.bcond_periodic_gpu(*args)
        elif c.boundary == "channel":
            L.bcond_channel_gpu(*args)
        else:
            W = self.R
            L.bcond_inflow_gpu(*args, R(c.u_inflow), R(3.0), R(1.5), R(3.0), R(W(1.0 / 18.0)), R(W(1.0 / 36.0)),
                               R(1.0 - self.cte1))
        if self.obs_buf is not None:
            L.bcond_obs_gpu(self.pa, self.obs_buf.ptr, c.l, c.m, c.n, self.nx, self.ny, *self.box, c.l, c.m)

    def collision(self):
        c, L, R, W = self.case, self.lib, self.Rc, self.R
        if c.les:
            L.col_mc_les_gpu(self.pa, self.pb, c.l, c.m, c.n, self.nx, self.ny, R(self.omega), R(c.cte_s),
                             R(self.fgrad), c.force_axis if c.boundary == "channel" else 0)
        elif c.boundary == "channel":
            L.col_mc_force_gpu(self.pa, self.pb, c.l, c.m, c.n, self.nx, self.ny, R(self.omega), R(self.fgrad),
                               c.force_axis)
        else:
            L.col_mc_gpu(*self.pa, *list(self.pb)[:18], c.l, c.m, c.n, self.nx, self.ny, R(self.omega),
                         R(1.0 - self.cte1), R(self.cte1), R(W(1.0 / 3.0)), R(W(1.0 / 18.0)), R(W(1.0 / 36.0)),
                         R(3.0), R(1.5), R(3.0), R(0.0), R(0.0), R(0.0))
        for d in range(18):
            self.pa[d], self.pb[d] = self.pb[d], self.pa[d]

    def step(self, nsteps: int = 1):
        for _ in range(nsteps):
            self.boundaries()
            self.collision()
            self.itime += 1

    def sync(self):
        self.lib.gpu_device_sync()

    # ---- diagnostics (device-side) ----------------------------------------
    def diagno(self) -> np.ndarray:
        """{<u>, <v>, <w>, <rho>, <ke>} (SRC/diagno.F90)."""
        c = self.case
        out = np.zeros(5)
        rt.check(self.lib.diagno_gpu(self.pa, c.l, c.m, c.n, self.nx, self.ny, self.Rc(self.cte1),
                                     out.ctypes.data), "diagno_gpu")
        return out

    def varm(self):
        c = self.case
        u, w, v = (np.zeros(c.n) for _ in range(3))
        den = np.zeros(c.m)
        rt.check(self.lib.varm_gpu(self.pa, c.l, c.m, c.n, self.nx, self.ny, self.Rc(self.cte1), u.ctypes.data,
                                   w.ctypes.data, v.ctypes.data, den.ctypes.data), "varm_gpu")
        return u, w, v, den

    def prof_k(self, ic=None, jc=None) -> np.ndarray:
        c = self.case
        ic = c.l // 2 if ic is None else ic
        jc = c.m // 2 if jc is None else jc
        out = np.zeros((c.n, 4), dtype=self.R)
        rt.check(self.lib.prof_k_gpu(self.pa, c.l, c.m, c.n, self.nx, self.ny, ic, jc, self.Rc(self.cte1),
                                     out.ctypes.data), "prof_k_gpu")
        return out

    def prof_i(self, jc=None, kc=None) -> np.ndarray:
        c = self.case
        jc = c.m // 2 if jc is None else jc
        kc = c.n // 2 if kc is None else kc
        out = np.zeros((c.l, 4), dtype=self.R)
        rt.check(self.lib.prof_i_gpu(self.pa, c.l, c.m, c.n, self.nx, self.ny, jc, kc, self.Rc(self.cte1),
                                     out.ctypes.data), "prof_i_gpu")
        return out

    def prof_j(self, ic=None, kc=None) -> np.ndarray:
        c = self.case
        ic = c.l // 2 if ic is None else ic
        kc = c.n // 2 if kc is None else kc
        out = np.zeros((c.m, 4), dtype=self.R)
        rt.check(self.lib.prof_j_gpu(self.pa, c.l, c.m, c.n, self.nx, self.ny, ic, kc, self.Rc(self.cte1),
                                     out.ctypes.data), "prof_j_gpu")
        return out

    def probe(self, i0, j0, k0) -> np.ndarray:
        c = self.case
        out = np.zeros(4, dtype=self.R)
        rt.check(self.lib.probe_gpu(self.pa, c.l, c.m, c.n, self.nx, self.ny, i0, j0, k0, self.Rc(self.cte1),
                                    out.ctypes.data), "probe_gpu")
        return out

    def draglift(self) -> np.ndarray:
        c = self.case
        if self.obs_buf is None:
            raise rt.BGKError("draglift needs an obstacle")
        out = np.zeros(3)
        rt.check(self.lib.draglift_gpu(self.pa, self.obs_buf.ptr, c.l, c.m, c.n, self.nx, self.ny, *self.centre,
                                       self.radius, self.Rc(c.u_inflow), out.ctypes.data), "draglift_gpu")
        return out

    def close(self):
        for buf in self._a + self._b:
            buf.close()
        if self.obs_buf is not None:
            self.obs_buf.close()
