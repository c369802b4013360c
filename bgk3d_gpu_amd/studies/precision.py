"""Storage-precision study on the body-force Poiseuille channel.

Configuration: 512 x 128 x 128 cells, periodic in x and z, half-way
bounce-back walls at y = 1/2 and m + 1/2 (`--walls z`: the reference's
orientation, walls at z = 1/2, n + 1/2 and y periodic,
SRC/bcond.channel.F90:255-267; on this square cross section the two agree
after exchanging y and z), body force g_x = 1e-6 per step (10 * fgrad,
fgrad = 4 u0 nu / (5 m^2), SRC/hencol.F90:52, so u0 = 0.02048), tau = 0.8
(nu = 0.1), rest start, 50 000 steps.

For each storage form (FP64, FP32, FP16, BF16; 16-bit with nearest and
stochastic rounding) it reports MLUPS and the relative L2 error of the
velocity field against the FP64 run (the FP64 GPU oracle) and of the
plane-averaged u_x across the channel against the analytic start-up solution of the
discrete problem (steady parabola of curvature -g/nu plus the decaying
Fourier series, plus the +g/2 the post-collision momentum carries).

    python -m bgk3d_gpu_amd.studies.precision [--steps N] [--size L M N] [--walls y|z] > out.json
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from ..cases import Case
from ..engine import Engine


def analytic_ux(n, g, nu, t, terms=200):
    """u_x across a channel of n cells (walls at 1/2 and n + 1/2) at step t,
    start from rest; the stored (post-collision) momentum adds g/2."""
    H = float(n)
    z = np.arange(1, n + 1) - 0.5          # distance from the lower wall
    u = g / (2 * nu) * z * (H - z)
    for q in range(terms):
        kq = (2 * q + 1) * np.pi / H
        u -= 4 * g * H ** 2 / (nu * np.pi ** 3 * (2 * q + 1) ** 3) * np.sin(kq * z) * np.exp(-nu * kq ** 2 * t)
    return u + g / 2


def run(storage, rounding, case, steps, chunk=5000):
    with Engine(case, storage, rounding=rounding, seed=12345) as eng:
        t0 = time.perf_counter()
        total_ms = 0.0
        done = 0
        while done < steps:
            s = min(chunk, steps - done)
            ms, _ = eng.time_steps(s)
            total_ms += ms
            done += s
        wall = time.perf_counter() - t0
        rho, ux, uy, uz = eng.fields()
    return {"mlups": case.cells * steps / (total_ms * 1e-3) / 1e6, "wall_s": wall}, (ux, uy, uz)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=50000)
    ap.add_argument("--size", type=int, nargs=3, default=[512, 128, 128])
    ap.add_argument("--walls", choices=["y", "z"], default="y")
    args = ap.parse_args(argv)
    l, m, n = args.size
    nu, g = 0.1, 1e-6
    u0 = (g / 10) * 5 * m * m / (4 * nu)
    case = Case("channel", l, m, n, svisc=nu, u0=u0, wall_axis=2 if args.walls == "y" else 3)
    width = m if args.walls == "y" else n
    avg_axes = (0, 2) if args.walls == "y" else (1, 2)  # fields are (k, j, i)
    forms = [("f64", 0), ("f32", 0), ("f16", 0), ("f16", 1), ("bf16", 0), ("bf16", 1)]
    results = {}
    ref = None
    exact = analytic_ux(width, g, nu, args.steps)
    for storage, rnd in forms:
        name = storage + ("_sr" if rnd else "")
        print(f"[precision] {name} ...", file=sys.stderr, flush=True)
        info, (ux, uy, uz) = run(storage, rnd, case, args.steps)
        if ref is None:
            ref = (ux, uy, uz)
        num = np.sqrt(((ux - ref[0]) ** 2 + (uy - ref[1]) ** 2 + (uz - ref[2]) ** 2).sum())
        den = np.sqrt((ref[0] ** 2 + ref[1] ** 2 + ref[2] ** 2).sum())
        prof = ux.mean(axis=avg_axes)
        info.update({
            "l2_rel_vs_f64": float(num / den),
            "l2_rel_profile_vs_analytic": float(np.sqrt(((prof - exact) ** 2).sum() / (exact ** 2).sum())),
            "u_center": float(prof[width // 2]),
        })
        results[name] = info
        print(f"[precision] {name}: {info}", file=sys.stderr, flush=True)
    out = {"study": "poiseuille_precision", "grid": [l, m, n], "walls": args.walls, "steps": args.steps,
           "tau": 0.8, "g": g, "u_center_analytic": float(exact[width // 2]), "forms": results}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
