"""Rate of the reference-compatible ABI by row padding: periodic Taylor-Green,
the lid-driven cavity (--boundary ldc) or the flow past the sphere (--boundary
inflow: inflow/outflow in x, the reference's obstacle shell).

Drives libbgk3d_gpu_{sp,dp}.so the reference's way (bcond_periodic_gpu,
bcond_driven_gpu or bcond_inflow_gpu + bcond_obs_gpu, then col_mc_gpu, then
the swap, every step; caller-owned arrays of
(0:l+1+ipad, 0:m+1, 0:n+1)) and prints one JSON line per precision and
padding: the mean step time from hipEvents around the timed steps, and
MLUPS.  With ipad = 0 the rows hold l + 2 values (258 at the default size)
and start anywhere in a 128-byte line; larger paddings show how the launcher's choice between
row-by-row waves and waves over each plane's interior rows plays out.

    python -m bgk3d_gpu_amd.studies.reference_abi > reference_abi.json
    python -m bgk3d_gpu_amd.studies.reference_abi --prec sp --ipad 0 --steps 20
    python -m bgk3d_gpu_amd.studies.reference_abi --boundary ldc --size 64 --ipad 0
    python -m bgk3d_gpu_amd.studies.reference_abi --boundary inflow --grid 1024,384,384 --ipad 0 --prec sp
"""
from __future__ import annotations

import argparse
import json

from .. import runtime as rt
from ..cases import Case
from ..simulation import Simulation


def time_steps(case: Case, prec: str, steps: int, warmup: int) -> float:
    sim = Simulation(case, prec, poison=False, device_init=True)
    try:
        sim.step(warmup)
        sim.sync()
        start, stop = rt.Event(sim.lib), rt.Event(sim.lib)
        start.record()
        sim.step(steps)
        stop.record()
        return start.elapsed_ms(stop)
    finally:
        sim.close()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--grid", default=None, help="l,m,n (default: size,size,size)")
    ap.add_argument("--boundary", choices=["periodic", "ldc", "inflow"], default="periodic")
    ap.add_argument("--prec", default="sp,dp", help="comma-separated: sp, dp")
    ap.add_argument("--ipad", default="0,1,30,70", help="comma-separated row paddings")
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    args = ap.parse_args(argv)
    l, m, n = (int(x) for x in args.grid.split(",")) if args.grid else (args.size,) * 3
    for prec in args.prec.split(","):
        for ipad in (int(x) for x in args.ipad.split(",")):
            if args.boundary == "periodic":
                case = Case("periodic", l, m, n, svisc=0.01, u0=0.05, init=1, ipad=ipad)
            elif args.boundary == "ldc":  # Re = 1000 with the reference's lid speed
                case = Case("ldc", l, m, n, svisc=0.1 * l / 1000.0, u0=0.1, ipad=ipad)
            else:  # the reference's sphere shell, u_inflow = 0.1
                case = Case("inflow", l, m, n, svisc=0.02, u0=0.1, obstacle=True, ipad=ipad)
            ms = time_steps(case, prec, args.steps, args.warmup)
            print(json.dumps({"boundary": args.boundary, "grid": [l, m, n], "prec": prec, "ipad": ipad,
                              "nx": l + 2 + ipad, "step_us": round(ms / args.steps * 1e3, 1),
                              "mlups": round(case.cells * args.steps / (ms * 1e-3) / 1e6, 1)}), flush=True)


if __name__ == "__main__":
    main()
