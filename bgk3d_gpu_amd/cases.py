"""Run configurations of the reference, expressed at run time.

The reference selects the boundary family and test case with compile-time
flags (SRC/Makefile:327-362: LDC, POF = CHANNEL, TGV/HIT = PERIODIC(+TGV3D),
VKS = INFLOW + OBSTACLE + DRAG) and reads sizes and physics from the
`&parameters` namelist of bgk.input (SRC/input.F90:44-62).  A `Case` holds
both; `read_namelist` accepts the reference's bgk.input files unchanged.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field, replace
from pathlib import Path

import numpy as np

BOUNDARIES = ("ldc", "periodic", "channel", "inflow")

# u_inflow is hard-wired as the default-real literal 0.1 (SRC/bcond.inflow.F90:81)
U_INFLOW_REF = float(np.float32(0.1))
# namelist cteS default 0.1, a default-real literal (SRC/input.F90:59)
CTE_S_REF = float(np.float32(0.1))


@dataclass
class Case:
    boundary: str = "ldc"          # ldc | periodic | channel | inflow
    l: int = 64
    m: int = 64
    n: int = 64
    svisc: float = 0.064           # kinematic viscosity (lattice units)
    u0: float = 0.1                # lid velocity, or volume-force scale for channel
    init: int = 0                  # 0 rest, 1 3-D Taylor-Green (periodic)
    ipad: int = 0
    jpad: int = 0
    obstacle: bool = False         # sphere shell (inflow case)
    u_inflow: float = U_INFLOW_REF
    force_axis: int = 1            # channel body-force axis (1 x, 2 y, 3 z)
    les: bool = False              # Smagorinsky LES relaxation (the reference's -DLES builds)
    cte_s: float = CTE_S_REF       # Smagorinsky constant (namelist cteS)
    wall_axis: int = 3             # channel walls: 3 at z = 1, n (the reference), 2 at y = 1, m
    noshift: bool = False          # unshifted storage, the reference's -DNOSHIFT builds (reference ABI only)
    extra: dict = field(default_factory=dict)

    def __post_init__(self):
        if self.boundary not in BOUNDARIES:
            raise ValueError(f"unknown boundary family {self.boundary!r}")
        if min(self.l, self.m, self.n) < 1:
            raise ValueError("grid sizes must be positive")

    # ---- parameters derived as the reference derives them ----------------
    @property
    def u00(self) -> float:
        """Lid velocity after build_bcond (SRC/build.bcond.F90:29-35)."""
        return 0.0 if self.boundary == "channel" else self.u0

    @property
    def u_force(self) -> float:
        """Volume-force velocity scale after build_bcond."""
        return self.u0 if self.boundary == "channel" else 0.0

    def hencol(self, prec: str = "sp"):
        """omega = 2/(6 nu + 1), fgrad = 4 u0 nu / (5 m^2) in the working kind
        (SRC/hencol.F90:31, :52)."""
        R = np.float32 if prec == "sp" else np.float64
        nu, u0 = R(self.svisc), R(self.u_force)
        omega = (R(1) + R(1)) / ((R(6.0) * nu) + R(1.0))
        fgrad = R(4.0) * u0 * nu / (R(5.0) * R(self.m) * R(self.m))
        return omega, fgrad

    def lid_force(self, prec: str = "sp"):
        """force = u00/(6.0) (SRC/bcond.driven.F90:61)."""
        R = np.float32 if prec == "sp" else np.float64
        return R(self.u00) / R(6.0)

    @property
    def shape(self):
        """(nz, ny, nx) of one population array incl. halo and padding."""
        return (self.n + 2, self.m + 2 + self.jpad, self.l + 2 + self.ipad)

    @property
    def cells(self) -> int:
        return self.l * self.m * self.n

    def with_(self, **kw) -> "Case":
        return replace(self, **kw)


def sphere_obstacle(l: int, m: int, n: int, centre=None, radius=None):
    """Hollow sphere shell of build_obs (SRC/build.obs.F90:34-97):
    (R-2)^2 < d^2 < (R+2)^2 around (2l/5, m/2, n/2), R = 32 if m > 256 else
    m/8.  Returns (obs int32 (n, m, l) = Fortran (1:l,1:m,1:n), box
    (imin, imax, jmin, jmax, kmin, kmax), radius, centre)."""
    ic, jc, kc = centre if centre is not None else (2 * l // 5, m // 2, n // 2)
    rad = radius if radius is not None else (32 if m > 256 else m // 8)
    k, j, i = np.meshgrid(np.arange(1, n + 1), np.arange(1, m + 1), np.arange(1, l + 1), indexing="ij")
    d2 = (ic - i) ** 2 + (jc - j) ** 2 + (kc - k) ** 2
    obs = ((d2 > (rad - 2) ** 2) & (d2 < (rad + 2) ** 2)).astype(np.int32)
    if obs.any():
        kk, jj, ii = np.nonzero(obs)
        box = (int(ii.min()) + 1, int(ii.max()) + 1, int(jj.min()) + 1, int(jj.max()) + 1, int(kk.min()) + 1,
               int(kk.max()) + 1)
    else:
        box = (l, 0, m, 0, n, 0)
    return obs, box, rad, (ic, jc, kc)


_NML_KEYS = {"svisc", "u0", "itfin", "ivtim", "isignal", "itsave", "icheck", "irestart", "init_v", "lx", "ly",
             "lz", "flag1", "flag2", "flag3", "ipad", "jpad", "radius", "ctes"}


def read_namelist(path) -> dict:
    """Parse the `&parameters` namelist of a reference bgk.input file."""
    text = Path(path).read_text()
    m = re.search(r"&parameters(.*?)/", text, re.S | re.I)
    if not m:
        raise ValueError(f"{path}: no &parameters namelist")
    out = {}
    for key, val in re.findall(r"(\w+)\s*=\s*([^,\s/]+)", m.group(1)):
        key = key.lower()
        if key not in _NML_KEYS:
            raise ValueError(f"{path}: unknown namelist key {key}")
        v = val.lower().replace("d", "e")
        out[key] = float(v) if any(c in v for c in ".e") else int(v)
    return out


def case_from_namelist(path, boundary: str = "ldc") -> tuple[Case, dict]:
    """Case + run controls (itfin, icheck, ...) from a reference bgk.input."""
    nml = read_namelist(path)
    case = Case(boundary=boundary, l=int(nml["lx"]), m=int(nml["ly"]), n=int(nml["lz"]), svisc=float(nml["svisc"]),
                u0=float(nml["u0"]), init=1 if boundary == "periodic" else 0, ipad=int(nml.get("ipad", 0)),
                jpad=int(nml.get("jpad", 0)), obstacle=(boundary == "inflow"))
    return case, nml
