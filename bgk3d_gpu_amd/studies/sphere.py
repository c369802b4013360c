"""Flow past a sphere: 1024 x 384 x 384 cells, Re = 300.

1024 x 384 x 384 cells (151M), FP32, solid sphere of radius 32 centred at
(256, 192, 192), inflow u = (0.05, 0, 0) at x = 0 and the reference's
equilibrium outflow at x = l+1 (SRC/bcond.inflow.F90:131-181), periodic in y
and z, Re = 300 on the diameter (nu = 0.05 * 64 / 300), rest start.  Drag
and lift by momentum exchange with FP64 accumulation every 100 steps
(SRC/draglift.F90), reported both in the reference's normalisation
F / (u_inflow^2 R) and as a drag coefficient C = F / (0.5 u^2 pi R^2).
Memory: two lattices of 19 population planes per z level, 23.9 GB.

    python -m bgk3d_gpu_amd.studies.sphere [--steps N] [--every K] > out.json
"""
from __future__ import annotations

import argparse
import json
import math
import sys

from ..cases import Case
from ..engine import Engine


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--size", type=int, nargs=3, default=[1024, 384, 384])
    ap.add_argument("--radius", type=int, default=32)
    ap.add_argument("--shell", action="store_true", help="the reference's hollow shell instead of a solid sphere")
    args = ap.parse_args(argv)
    l, m, n = args.size
    U, R, Re = 0.05, args.radius, 300.0
    nu = U * 2 * R / Re
    centre = (l // 4, m // 2, n // 2)
    case = Case("inflow", l, m, n, svisc=nu, u0=U, obstacle=True)
    series = []
    total_ms = 0.0
    with Engine(case, "f32", obstacle=1 if args.shell else 2, centre=centre, radius=R, u_inflow=U) as eng:
        info = eng.info()
        print(f"[sphere] {l}x{m}x{n}, {info['bytes'] / 1e9:.1f} GB, nu={nu:.6f}", file=sys.stderr, flush=True)
        done = 0
        while done < args.steps:
            s = min(args.every, args.steps - done)
            ms, _ = eng.time_steps(s)
            total_ms += ms
            done += s
            fx, fy, fz = eng.draglift()
            series.append([done, fx, fy, fz])
            if done % (10 * args.every) == 0:
                print(f"[sphere] step {done}: F/(u^2 R) = {fx:.5f} {fy:.2e} {fz:.2e}", file=sys.stderr, flush=True)
        diag = eng.diagno()
    # F/(u^2 R) -> C_D = F / (0.5 u^2 pi R^2) = (F/(u^2 R)) * 2 / (pi R)
    to_cd = 2.0 / (math.pi * R)
    tail = [r[1] for r in series[-max(1, len(series) // 5):]]
    out = {"study": "sphere", "grid": [l, m, n], "cells": l * m * n, "steps": args.steps, "radius": R,
           "centre": centre, "u_inflow": U, "Re": Re, "nu": nu, "obstacle": "shell" if args.shell else "solid",
           "mlups": case.cells * args.steps / (total_ms * 1e-3) / 1e6, "device_bytes": info["bytes"],
           "cd_last20pct_mean": sum(tail) / len(tail) * to_cd, "diagno": list(diag),
           "series_F_over_u2R": series}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
