#!/usr/bin/env python3
"""Generate golden fixtures from the reference's own Fortran CPU routines.

Runs oracle/_ref/<case>_<prec>/ref_harness.x (the reference sources compiled
by oracle/Makefile, linked with our small harness) and records, per run:
  * sha256 of the interior populations (fluid cells only when there is an
    obstacle), C order (k, j, i), native little-endian bytes;
  * per-population sums in float64 (to locate a mismatch);
  * the reference's own text outputs (diagno.dat, probe.dat, prof_i.dat,
    prof_j.dat, prof_k.dat, u_med.dat, drag.lift.dat) as written by its
    diagnostics;
  * digests of the files its writers produce at the last step: the restart
    file of save_raw (time step and the interior of every record -- the halo
    cells hold stale boundary values nobody reads), save_vel, and the VTK
    planes vtk_xy/xz/yz_bin and vtk_om_xy_bin (whole files);
  * a restart pin: the reference's own save_raw at half the steps, restored
    by its restore_raw and run on, ends on the full run's state;
  * LES runs from the reference's -DLES builds, and the ORIGINAL two-pass
    build's final state next to the fused build's;
  * for the single-routine pins: a seeded random state pushed through one
    `boundaries` call or one `col_MC` call.

Output: tests/golden/fixtures.json.  The reference itself is never stored.
Usage: python tests/golden/make_fixtures.py   (needs oracle/_ref built)
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import format_support as F  # noqa: E402
import oracle_support as O  # noqa: E402

# (name, case, l, m, n, steps, svisc, u0, init_v, icheck)
RUNS = [
    ("ldc_24", "ldc", 24, 24, 24, 300, 0.064, 0.1, 0, 100),
    ("ldc_64_re100", "ldc", 64, 64, 64, 2000, 0.064, 0.1, 0, 1000),
    ("tgv_32", "periodic", 32, 32, 32, 200, 0.01, 0.05, 1, 100),
    ("channel_64x32x32", "channel", 64, 32, 32, 500, 0.1, 0.02048, 0, 250),
    ("inflow_sphere_100x48x48", "inflow", 100, 48, 48, 300, 0.02, 0.1, 0, 100),
]

# LES builds (-DLES, SRC/col_MC.F90:381-466; cteS = 0.1): low viscosities
# where the Smagorinsky rate departs from omega
RUNS_LES = [
    ("ldc_20x18x16_les", "ldc", 20, 18, 16, 200, 0.002, 0.1, 0, 100),
    ("tgv_24x20x16_les", "periodic", 24, 20, 16, 100, 0.001, 0.05, 1, 50),
    ("channel_24x16x20_les", "channel", 24, 16, 20, 200, 0.01, 0.02, 0, 100),
]

# the reference's mixed-precision builds (SRC/precision.F90:30-43, Makefile
# flags -DHALF_P -DMIXEDPRECISION and -DMIXEDPRECISION): m1 stores real*2
# and computes in single precision, m2 stores single and computes in double.
# MIXED1 sets the volume force to zero (SRC/hencol.F90:35-47), so its
# channel would stay at rest: cavity and Taylor-Green only.
RUNS_MIXED = [
    ("ldc_24", "ldc", 24, 24, 24, 300, 0.064, 0.1, 0, 100, ("m1", "m2")),
    ("tgv_32", "periodic", 32, 32, 32, 200, 0.01, 0.05, 1, 100, ("m1", "m2")),
    ("ldc_64_re100", "ldc", 64, 64, 64, 2000, 0.064, 0.1, 0, 1000, ("m2",)),
    ("channel_64x32x32", "channel", 64, 32, 32, 500, 0.1, 0.02048, 0, 250, ("m2",)),
]

# unshifted storage (-DNOSHIFT: f itself instead of f - w,
# SRC/col_MC.F90:139-144, SRC/init.F90:71-75)
RUNS_NOSHIFT = [
    ("ldc_24", "ldc", 24, 24, 24, 300, 0.064, 0.1, 0, 100),
    ("tgv_32", "periodic", 32, 32, 32, 200, 0.01, 0.05, 1, 100),
    ("channel_64x32x32", "channel", 64, 32, 32, 500, 0.1, 0.02048, 0, 250),
    ("inflow_sphere_100x48x48", "inflow", 100, 48, 48, 300, 0.02, 0.1, 0, 100),
]

# single-routine pins on seeded random states: (name, case, l, m, n, mode)
UNIT = [
    ("bc_driven", "ldc", 13, 11, 9, 1),
    ("bc_periodic", "periodic", 13, 11, 9, 1),
    ("bc_channel", "channel", 13, 11, 9, 1),
    ("bc_inflow_obs", "inflow", 40, 24, 24, 1),
    ("col_mc", "ldc", 13, 11, 9, 2),
    ("col_mc_channel_force", "channel", 13, 11, 9, 2),
]


def interior_bytes(A, obs=None):
    I = np.ascontiguousarray(A[:, 1:-1, 1:-1, 1:-1])
    if obs is not None:
        I = np.ascontiguousarray(I[:, obs == 0])
    return I


def digest(arr) -> str:
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def random_state(l, m, n, prec, seed):
    rng = np.random.default_rng(seed)
    A = O.new_lattice(l, m, n, prec)
    A[...] = (rng.standard_normal(A.shape) * 0.01).astype(A.dtype)
    return A


def main():
    out = {"generator": "tests/golden/make_fixtures.py", "runs": {}, "unit": {}, "restart": {}, "original": {},
           "mixed": {}, "noshift": {}}
    for name, case, l, m, n, steps, svisc, u0, init_v, icheck in RUNS:
        for prec in ("sp", "dp"):
            if not O.ref_available(case, prec):
                raise SystemExit(f"oracle/_ref/{case}_{prec} not built")
            A, texts, stdout = O.run_ref(case, l, m, n, steps, prec=prec, svisc=svisc, u0=u0, icheck=icheck,
                                         init_v=init_v, write=True)
            obs = texts.get("obs")
            I = interior_bytes(A, obs)
            out["runs"][f"{name}_{prec}"] = {
                "case": case, "l": l, "m": m, "n": n, "steps": steps, "svisc": svisc, "u0": u0, "init": init_v,
                "icheck": icheck, "prec": prec, "sha256_interior": digest(I),
                # every interior cell, solid cells included (the serial
                # loop order of SRC/bcond.obs.F90:101-122 fixes them)
                "sha256_interior_all": digest(interior_bytes(A)),
                "pop_sums": [float(x) for x in I.reshape(19, -1).astype(np.float64).sum(axis=1)],
                "texts": {k: v for k, v in texts.items() if k not in ("obs", "files")},
                "files": F.file_digests(texts["files"], l, m, n, prec),
            }
            print(name, prec, "done", flush=True)
    # LES runs
    for name, case, l, m, n, steps, svisc, u0, init_v, icheck in RUNS_LES:
        for prec in ("sp", "dp"):
            if not O.ref_available(case, prec, variant="_les"):
                raise SystemExit(f"oracle/_ref/{case}_{prec}_les not built")
            A, texts, _ = O.run_ref(case, l, m, n, steps, prec=prec, svisc=svisc, u0=u0, icheck=icheck,
                                    init_v=init_v, variant="_les")
            I = interior_bytes(A)
            out["runs"][f"{name}_{prec}"] = {
                "case": case, "l": l, "m": m, "n": n, "steps": steps, "svisc": svisc, "u0": u0, "init": init_v,
                "icheck": icheck, "prec": prec, "les": True, "sha256_interior": digest(I),
                "pop_sums": [float(x) for x in I.reshape(19, -1).astype(np.float64).sum(axis=1)],
                "texts": {k: v for k, v in texts.items() if k not in ("obs", "files")},
            }
            print(name, prec, "done", flush=True)
    # mixed-precision and unshifted builds: whole interior, text outputs
    for name, case, l, m, n, steps, svisc, u0, init_v, icheck, precs in RUNS_MIXED:
        for prec in precs:
            if not O.ref_available(case, prec):
                raise SystemExit(f"oracle/_ref/{case}_{prec} not built")
            A, texts, _ = O.run_ref(case, l, m, n, steps, prec=prec, svisc=svisc, u0=u0, icheck=icheck,
                                    init_v=init_v)
            I = interior_bytes(A)
            out["mixed"][f"{name}_{prec}"] = {
                "case": case, "l": l, "m": m, "n": n, "steps": steps, "svisc": svisc, "u0": u0, "init": init_v,
                "icheck": icheck, "prec": prec, "sha256_interior": digest(I),
                "pop_sums": [float(x) for x in I.reshape(19, -1).astype(np.float64).sum(axis=1)],
                "texts": {k: v for k, v in texts.items() if k not in ("obs", "files")},
            }
            print("mixed", name, prec, "done", flush=True)
    for name, case, l, m, n, steps, svisc, u0, init_v, icheck in RUNS_NOSHIFT:
        for prec in ("sp", "dp"):
            if not O.ref_available(case, prec, variant="_noshift"):
                raise SystemExit(f"oracle/_ref/{case}_{prec}_noshift not built")
            A, texts, _ = O.run_ref(case, l, m, n, steps, prec=prec, svisc=svisc, u0=u0, icheck=icheck,
                                    init_v=init_v, variant="_noshift")
            I = interior_bytes(A)
            out["noshift"][f"{name}_{prec}"] = {
                "case": case, "l": l, "m": m, "n": n, "steps": steps, "svisc": svisc, "u0": u0, "init": init_v,
                "icheck": icheck, "prec": prec, "sha256_interior": digest(I),
                "pop_sums": [float(x) for x in I.reshape(19, -1).astype(np.float64).sum(axis=1)],
                "texts": {k: v for k, v in texts.items() if k not in ("obs", "files")},
            }
            print("noshift", name, prec, "done", flush=True)
    # the ORIGINAL two-pass build (movef + col) ends on the fused build's state
    for name, case, l, m, n, steps, svisc, u0, init_v, icheck in RUNS[:1] + RUNS[2:4]:
        for prec in ("sp", "dp"):
            A, _, _ = O.run_ref(case, l, m, n, steps, prec=prec, svisc=svisc, u0=u0, icheck=icheck, init_v=init_v,
                                variant="_orig")
            out["original"][f"{name}_{prec}"] = {
                "sha256_interior": digest(interior_bytes(A)),
                "equals_fused": digest(interior_bytes(A)) == out["runs"][f"{name}_{prec}"]["sha256_interior"]}
            print("original", name, prec, out["original"][f"{name}_{prec}"]["equals_fused"], flush=True)
    # restart pin: save_raw at half the steps, restore_raw, run on
    for name, case, l, m, n, steps, svisc, u0, init_v, icheck in RUNS[:1] + RUNS[2:3]:
        for prec in ("sp", "dp"):
            half = steps // 2
            _, texts, _ = O.run_ref(case, l, m, n, half, prec=prec, svisc=svisc, u0=u0, init_v=init_v, write=True)
            raw = texts["files"][f"save.{half:08d}.bin"]
            A, _, _ = O.run_ref(case, l, m, n, steps, prec=prec, svisc=svisc, u0=u0, restore=raw)
            full = out["runs"][f"{name}_{prec}"]
            ok = digest(interior_bytes(A)) == full["sha256_interior"]
            if not ok:
                raise SystemExit(f"restart of {name} {prec} does not reproduce the full run")
            out["restart"][f"{name}_{prec}"] = {
                "run": f"{name}_{prec}", "save_step": half,
                "save_raw": F.file_digests({f"save.{half:08d}.bin": raw}, l, m, n, prec)["save_raw"],
            }
            print("restart", name, prec, "done", flush=True)
    for name, case, l, m, n, mode in UNIT:
        for prec in ("sp", "dp"):
            seed = int(hashlib.sha256(f"{name}-{prec}".encode()).hexdigest()[:8], 16)
            A_in = random_state(l, m, n, prec, seed)
            A, texts, _ = O.run_ref(case, l, m, n, 1, prec=prec, svisc=0.02 if case != "channel" else 0.1,
                                    u0=0.1 if case != "channel" else 0.02048, mode=mode, A_in=A_in)
            # the boundary pins compare whole arrays (halo included); col_MC
            # writes the interior only (b is the caller's concern)
            out["unit"][f"{name}_{prec}"] = {
                "case": case, "l": l, "m": m, "n": n, "mode": mode, "prec": prec, "seed": seed,
                "svisc": 0.02 if case != "channel" else 0.1, "u0": 0.1 if case != "channel" else 0.02048,
                "sha256_full": digest(A), "sha256_interior": digest(interior_bytes(A)),
                "sha256_pulled": digest(O.pulled_set(A, l, m, n)),
            }
            print(name, prec, "done", flush=True)
    (HERE / "fixtures.json").write_text(json.dumps(out, indent=1, sort_keys=True))
    print("wrote", HERE / "fixtures.json")


if __name__ == "__main__":
    main()
