"""Fused-kernel rate against the lattice footprint.

Periodic Taylor-Green in FP32 (handle API) at sizes from 0.7 GB to 22 GB of
device memory: the same kernel and row length, only the number of rows and
planes grows.  Prints one JSON line per size: bytes held, mean step and
collide-kernel durations (hipEvents), MLUPS and the kernel's algorithmic
bandwidth (38 x 4 bytes per cell).

    python -m bgk3d_gpu_amd.studies.footprint > footprint.json
"""
from __future__ import annotations

import argparse
import json

from ..cases import Case
from ..engine import Engine

SIZES = [(256, 128, 128), (256, 256, 128), (256, 256, 256), (256, 256, 512), (256, 512, 512), (512, 512, 512)]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=60)
    args = ap.parse_args(argv)
    for l, m, n in SIZES:
        case = Case("periodic", l, m, n, svisc=0.01, u0=0.05, init=1)
        with Engine(case, "f32") as eng:
            info = eng.info()
            eng.time_steps(10)
            ms, kern_ms = eng.time_steps(args.steps)
        cells = l * m * n
        print(json.dumps({"grid": [l, m, n], "device_bytes": info["bytes"], "step_us": ms / args.steps * 1e3,
                          "collide_us": kern_ms * 1e3, "mlups": cells * args.steps / (ms * 1e-3) / 1e6,
                          "alg_GBs": 38 * 4 * cells / (kern_ms * 1e-3) / 1e9}), flush=True)


if __name__ == "__main__":
    main()
