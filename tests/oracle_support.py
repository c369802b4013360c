"""Test-side access to the CPU oracle and the compiled reference harness.

The oracle (oracle/_build/libbgk_oracle.so) is a C restatement of the
reference's Fortran CPU path; the reference harness binaries
(oracle/_ref/<case>_<prec>/ref_harness.x) are the reference's own routines
compiled from its sources.  Both are checkers: nothing under
``bgk3d_gpu_amd`` imports this module.

Arrays follow the reference layout: population d (0-based) of a lattice
is ``A[d]`` with numpy shape (nz, ny, nx) -- the Fortran column-major array
(0:nx-1, 0:ny-1, 0:nz-1) seen in C order (SRC/alloca.F90:153-191).
"""
from __future__ import annotations

import ctypes
import os
import shutil
import subprocess
import tempfile
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[1]
ORACLE_SO = REPO / "oracle" / "_build" / "libbgk_oracle.so"
REF_DIR = REPO / "oracle" / "_ref"

# u_inflow is hard-wired as the default-real literal 0.1 in the reference
# (SRC/bcond.inflow.F90:81), so the DP build runs with (double)0.1f.
U_INFLOW_REF = float(np.float32(0.1))

# boundary family codes shared with oracle_run_*: driven, periodic, channel,
# inflow (+ obstacle)
BC_CODE = {"ldc": 0, "periodic": 1, "channel": 2, "inflow": 3}

_lib = None


def oracle_lib():
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            subprocess.run(["make", "-C", str(REPO / "oracle"), "oracle"], check=True,
                           stdout=subprocess.DEVNULL)
        _lib = ctypes.CDLL(str(ORACLE_SO))
    return _lib


class noshift:
    """Context in which the oracle's routines use unshifted storage, the
    reference's -DNOSHIFT builds (cte1 per routine as the reference sets it,
    see oracle/bgk_oracle.c)."""

    def __enter__(self):
        oracle_lib().oracle_set_noshift(1)
        return self

    def __exit__(self, *exc):
        oracle_lib().oracle_set_noshift(0)


def np_dtype(prec: str):
    """Storage dtype of a build: sp, dp, and the reference's mixed builds
    m1 (real*2 storage, single compute) and m2 (single storage, double
    compute), SRC/precision.F90:30-43."""
    return {"sp": np.float32, "dp": np.float64, "m1": np.float16, "m2": np.float32}[prec]


def ctype(prec: str):
    return ctypes.c_float if prec == "sp" else ctypes.c_double


def sfx(prec: str) -> str:
    return "f32" if prec == "sp" else "f64"


def new_lattice(l, m, n, prec="sp", ipad=0, jpad=0, poison=True):
    nx, ny, nz = l + 2 + ipad, m + 2 + jpad, n + 2
    dt = np_dtype(prec)
    fill = np.finfo(dt).max if poison else 0
    return np.full((19, nz, ny, nx), fill, dtype=dt)


def ptrs(A):
    """Host array of the 19 base pointers (the Fortran pa(1:19))."""
    P = (ctypes.c_void_p * 19)()
    for d in range(19):
        P[d] = A[d].ctypes.data
    return P


def fn(name, prec):
    return getattr(oracle_lib(), f"oracle_{name}_{sfx(prec)}")


def hencol(svisc, u0, m, prec="sp"):
    ct = ctype(prec)
    om, fg = ct(), ct()
    f = fn("hencol", prec)
    f.argtypes = [ct, ct, ctypes.c_int, ctypes.POINTER(ct), ctypes.POINTER(ct)]
    f(ct(svisc), ct(u0), m, ctypes.byref(om), ctypes.byref(fg))
    return om.value, fg.value


def init(A, l, m, n, prec="sp", mode=0, u00=0.0):
    nx, ny = A.shape[3], A.shape[2]
    ct = ctype(prec)
    f = fn("init", prec)
    f.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 6 + [ct]
    f(ptrs(A), l, m, n, nx, ny, mode, ct(u00))


def build_obs(l, m, n):
    obs = np.zeros((n, m, l), dtype=np.int32)
    box = (ctypes.c_int * 6)()
    rad = ctypes.c_int()
    f = oracle_lib().oracle_build_obs
    f.restype = ctypes.c_long
    f(obs.ctypes.data_as(ctypes.c_void_p), l, m, n, box, ctypes.byref(rad))
    return obs, tuple(box), rad.value


# the reference's default Smagorinsky constant: cteS = 0.1, a default-real
# literal (SRC/input.F90:59), so the DP build sees the float-rounded value
CTE_S_REF = float(np.float32(0.1))


def col_mc(A, B, l, m, n, omega, prec="sp", force_mode=0, fgrad=0.0, fxyz=(0.0, 0.0, 0.0), nthreads=1,
           les=False, cte_s=CTE_S_REF):
    nx, ny = A.shape[3], A.shape[2]
    ct = ctype(prec)
    f = fn("col_mc", prec)
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 5 + [ct, ctypes.c_int, ct, ct, ct, ct, ct,
                                                                           ctypes.c_int]
    f(ptrs(A), ptrs(B), l, m, n, nx, ny, ct(omega), force_mode | (4 if les else 0), ct(fgrad), ct(fxyz[0]),
      ct(fxyz[1]), ct(fxyz[2]), ct(cte_s), nthreads)


def swap(A, B):
    """Fortran pointer swap of populations 1..18 (SRC/col_MC.F90:539-596)."""
    tmp = A[:18].copy()
    A[:18] = B[:18]
    B[:18] = tmp


def bcond(case, A, l, m, n, prec="sp", u00=0.0, u_inflow=U_INFLOW_REF, obs=None, box=None):
    nx, ny = A.shape[3], A.shape[2]
    ct = ctype(prec)
    P = ptrs(A)
    if case == "ldc":
        f = fn("bcond_driven", prec)
        f.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5 + [ct]
        f(P, l, m, n, nx, ny, ct(u00))
    elif case == "periodic":
        fn("bcond_periodic", prec)(P, l, m, n, nx, ny)
    elif case == "channel":
        fn("bcond_channel", prec)(P, l, m, n, nx, ny)
    elif case == "inflow":
        f = fn("bcond_inflow", prec)
        f.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5 + [ct]
        f(P, l, m, n, nx, ny, ct(u_inflow))
    if obs is not None:
        f = fn("bcond_obs", prec)
        f(P, obs.ctypes.data_as(ctypes.c_void_p), l, m, n, nx, ny, *box)


def run(case, A, B, l, m, n, steps, prec="sp", svisc=0.064, u0=0.1, obs=None, box=None, u_inflow=U_INFLOW_REF,
        nthreads=1, les=False, cte_s=CTE_S_REF):
    """Reference time loop (SRC/bgk3d.F90:169-196) on host arrays A, B.

    Pointer swaps are done on the pointer tables; on return the current
    lattice is copied back into A (populations 1..18 may live in B)."""
    nx, ny = A.shape[3], A.shape[2]
    ct = ctype(prec)
    omega, fgrad = hencol(svisc, u0, m, prec)
    # build_bcond (SRC/build.bcond.F90:29-35): lid velocity vs volume force
    if case == "channel":
        u00, force_mode = 0.0, 1
    else:
        u00, force_mode, fgrad = u0, 0, 0.0
    PA, PB = ptrs(A), ptrs(B)
    f = fn("run", prec)
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 6 + [ct, ct, ctypes.c_int, ct, ct,
                                                                           ctypes.c_void_p, ctypes.c_void_p,
                                                                           ctypes.c_int, ct, ctypes.c_int]
    obs_p = obs.ctypes.data_as(ctypes.c_void_p) if obs is not None else None
    box_p = (ctypes.c_int * 6)(*box) if box is not None else None
    f(PA, PB, l, m, n, nx, ny, BC_CODE[case], ct(omega), ct(u00), force_mode | (4 if les else 0), ct(fgrad),
      ct(u_inflow), obs_p, box_p, steps, ct(cte_s), nthreads)
    # gather the current lattice back into A
    cur = [None] * 19
    addr = {A[d].ctypes.data: ("A", d) for d in range(19)}
    addr.update({B[d].ctypes.data: ("B", d) for d in range(19)})
    for d in range(19):
        src, sd = addr[PA[d]]
        cur[d] = (A if src == "A" else B)[sd].copy()
    for d in range(19):
        A[d] = cur[d]


def diagno(A, l, m, n, prec="sp"):
    nx, ny = A.shape[3], A.shape[2]
    out = (ctypes.c_double * 5)()
    fn("diagno", prec)(ptrs(A), l, m, n, nx, ny, out)
    return np.array(out[:])


def varm(A, l, m, n, prec="sp"):
    nx, ny = A.shape[3], A.shape[2]
    dt = np_dtype(prec)
    u, w, v = (np.zeros(n, dt) for _ in range(3))
    den = np.zeros(m, dt)
    fn("varm", prec)(ptrs(A), l, m, n, nx, ny, *(x.ctypes.data_as(ctypes.c_void_p) for x in (u, w, v, den)))
    return u, w, v, den


def prof_k(A, l, m, n, ic, jc, prec="sp"):
    nx, ny = A.shape[3], A.shape[2]
    out = np.zeros((n, 4), np_dtype(prec))
    fn("prof_k", prec)(ptrs(A), l, m, n, nx, ny, ic, jc, out.ctypes.data_as(ctypes.c_void_p))
    return out


def prof_line(A, l, m, n, axis, c1, c2, prec="sp"):
    """Profile along axis 0 (i at (j, k) = (c1, c2)), 1 (j at (i, k)) or 2 (k
    at (i, j)): rows of (u, w, v, den) (SRC/prof_i.F90, prof_j.F90, prof_k.F90)."""
    nx, ny = A.shape[3], A.shape[2]
    out = np.zeros(((l, m, n)[axis], 4), np_dtype(prec))
    fn("prof_line", prec)(ptrs(A), l, m, n, nx, ny, axis, c1, c2, out.ctypes.data_as(ctypes.c_void_p))
    return out


def probe(A, i0, j0, k0, prec="sp"):
    """(u, v, w, rho) at one cell (SRC/probe.F90)."""
    nx, ny = A.shape[3], A.shape[2]
    out = np.zeros(4, np_dtype(prec))
    fn("probe", prec)(ptrs(A), nx, ny, i0, j0, k0, out.ctypes.data_as(ctypes.c_void_p))
    return out


def draglift(A, obs, l, m, n, radius, u_inflow=U_INFLOW_REF, prec="sp"):
    nx, ny = A.shape[3], A.shape[2]
    ct = ctype(prec)
    out = np.zeros(3, np_dtype(prec))
    f = fn("draglift", prec)
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 6 + [ct, ctypes.c_void_p]
    f(ptrs(A), obs.ctypes.data_as(ctypes.c_void_p), l, m, n, nx, ny, radius, ct(u_inflow),
      out.ctypes.data_as(ctypes.c_void_p))
    return out


# ---------------------------------------------------------------------------
# compiled reference harness (only where oracle/_ref was built)
# ---------------------------------------------------------------------------

def ref_available(case="ldc", prec="sp", omp=False, variant="") -> bool:
    """variant: '' (fused build), '_orig' (ORIGINAL two-pass build), '_les'."""
    return (REF_DIR / f"{case}_{prec}{'_omp' if omp else variant}" / "ref_harness.x").exists()


def run_ref(case, l, m, n, steps, prec="sp", svisc=0.064, u0=0.1, icheck=None, init_v=0, ipad=0, jpad=0,
            mode=0, A_in=None, omp=False, threads=None, keep_dir=None, write=False, restore=None, variant=""):
    """Run the reference harness; returns (A_out, outdir_texts, stdout).

    write=True adds texts["files"] = {name: bytes} of what the reference's
    save_raw, save_vel and VTK writers produced at the last step;
    restore=<bytes> starts from that restart file (the reference's
    restore_raw) instead of init."""
    exe = REF_DIR / f"{case}_{prec}{'_omp' if omp else variant}" / "ref_harness.x"
    d = Path(keep_dir) if keep_dir else Path(tempfile.mkdtemp(prefix="bgkref_"))
    d.mkdir(parents=True, exist_ok=True)
    try:
        if A_in is not None:
            init_v = -1
            np.ascontiguousarray(A_in).tofile(d / "in.bin")
        if restore is not None:
            init_v = -2
            (d / "restore.bin").write_bytes(restore)
        env = dict(os.environ)
        if threads:
            env["OMP_NUM_THREADS"] = str(threads)
        cmd = [str(exe), str(d), str(l), str(m), str(n), repr(float(svisc)), repr(float(u0)), str(steps),
               str(icheck or max(steps, 1)), str(init_v), str(ipad), str(jpad), str(mode), "1" if write else "0"]
        out = subprocess.run(cmd, check=True, capture_output=True, text=True, env=env, cwd=d).stdout
        nx, ny, nz = l + 2 + ipad, m + 2 + jpad, n + 2
        A = np.fromfile(d / "out.bin", dtype=np_dtype(prec)).reshape(19, nz, ny, nx)
        texts = {}
        for name in ("diagno.dat", "prof_i.dat", "prof_j.dat", "prof_k.dat", "probe.dat", "u_med.dat",
                     "drag.lift.dat"):
            p = d / name
            texts[name] = p.read_text() if p.exists() else ""
        if write:
            texts["files"] = {p.name: p.read_bytes() for p in sorted(d.iterdir())
                              if p.name.startswith(("save.", "save_vel.", "tec_"))}
        if (d / "obs.bin").exists():
            texts["obs"] = np.fromfile(d / "obs.bin", dtype=np.int32).reshape(n, m, l)
        return A, texts, out
    finally:
        if keep_dir is None:
            shutil.rmtree(d, ignore_errors=True)


# ---------------------------------------------------------------------------
# Fortran-style formatting, to compare against the reference's text outputs
# ---------------------------------------------------------------------------

def fortran_e(x: float, w: int = 14, d: int = 6) -> str:
    """Fortran Ew.d edit descriptor: 0.dddddd E+xx, right-justified."""
    x = float(x)
    if x == 0.0:
        body = "0." + "0" * d + "E+00"
        return body.rjust(w)
    mant, exp = f"{abs(x):.{d - 1}e}".split("e")
    digits = mant.replace(".", "")
    e = int(exp) + 1
    body = f"{'-' if x < 0 else ''}0.{digits}E{'+' if e >= 0 else '-'}{abs(e):02d}"
    return body.rjust(w)


def diagno_line(itime: int, vals) -> str:
    """SRC/diagno.F90:145,159: (i8,5(e14.6,1x)) of itime, <u>, <v>, <w>, <rho>, <ke>."""
    return f"{itime:8d}" + "".join(fortran_e(v) + " " for v in vals)


def parse_diagno(text: str):
    """diagno.dat rows as [(itime, [5 floats]), ...]."""
    rows = []
    for line in text.splitlines():
        if line.strip():
            t, *v = line.split()
            rows.append((int(t), [float(x) for x in v]))
    return rows


def assert_diagno_close(ours, ref_text, atol=1e-12):
    """Our diagnostics [(itime, 5 values), ...] against a reference
    diagno.dat: the same time steps, and every value within half a unit of
    the file's last printed digit (e14.6: 6 significant digits) of the
    printed value, except means that are zero up to rounding (|value| below
    atol: the sum of antisymmetric velocities, which an FP64 sum in another
    order leaves at a different 1e-17-size residue)."""
    ref = parse_diagno(ref_text)
    assert [t for t, _ in ours] == [t for t, _ in ref]
    for (t, a), (_, b) in zip(ours, ref):
        for q, (x, y) in enumerate(zip(np.asarray(a, dtype=float), b)):
            if abs(y) < atol and abs(x) < atol:
                continue
            half_unit = 0.5 * 10.0 ** (np.floor(np.log10(abs(y))) + 1 - 6)
            assert abs(x - y) <= half_unit * (1 + 1e-9), f"itime {t} column {q}: {x!r} vs printed {y!r}"


def parse_blocks(text: str):
    """Split a prof_k.dat / u_med.dat text into {t: [[floats], ...]} blocks."""
    blocks, cur, t = {}, None, None
    for line in text.splitlines():
        if line.startswith("# t="):
            t = int(line.split("=")[1])
            cur = blocks.setdefault(t, [])
        elif line.strip() and cur is not None:
            cur.append(line)
    return blocks


CX = [1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0]
CY = [-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1, 0]
CZ = [0, -1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, 1, 0, -1, -1, 0, 1, 0]


def pulled_set(A, l, m, n):
    """Concatenation, over d, of the cells the next col_MC pull reads from
    population d: x - c_d for every interior x (SRC/col_MC.F90:275-293)."""
    parts = []
    for d in range(19):
        parts.append(np.ascontiguousarray(
            A[d, 1 - CZ[d]:n + 1 - CZ[d], 1 - CY[d]:m + 1 - CY[d], 1 - CX[d]:l + 1 - CX[d]]).ravel())
    return np.concatenate(parts)
