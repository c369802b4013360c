"""Build the gfx950 HIP libraries in-tree with hipcc.

Outputs (bgk3d_gpu_amd/lib/):
  libbgk3d_gpu_sp.so   reference-compatible C-ABI, real_t = float
  libbgk3d_gpu_dp.so   reference-compatible C-ABI, real_t = double
  libbgk3d_engine.so   handle API: library-owned lattices, fused boundaries,
                       FP16/BF16 storage, batched serving

Flags: -ffp-contract=off keeps every product and sum rounded separately, as
the reference's CPU build does (the step is bandwidth-bound, so FMA would buy
nothing); packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
INCLUDE = PKG.parent / "include"

ARCH = os.environ.get("BGK_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]

TARGETS = {
    "libbgk3d_gpu_sp.so": (["v1_abi.hip"], []),
    "libbgk3d_gpu_dp.so": (["v1_abi.hip"], ["-DBGK_DOUBLE"]),
}
ENGINE_SOURCES = ["engine.hip", "serving.hip"]
if (CSRC / "engine.hip").exists():
    TARGETS["libbgk3d_engine.so"] = (ENGINE_SOURCES, [])


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP libraries cannot be built")


def build_one(name: str, sources, extra, verbose=False) -> Path:
    LIBDIR.mkdir(parents=True, exist_ok=True)
    out = LIBDIR / name
    tmp = out.with_suffix(".so.tmp")
    cmd = [hipcc(), *COMMON, *extra, *[str(CSRC / s) for s in sources], "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(x for x in r.stderr.splitlines() if "packed-fp32-ops" not in x)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {name}:\n{errs}")
    if errs.strip() and verbose:
        print(errs, file=sys.stderr)
    os.replace(tmp, out)
    return out


def build_fortran(verbose=False):
    """Fortran interface module + host program (amdflang), linked against
    libbgk3d_gpu_{sp,dp}.so with an $ORIGIN-relative rpath."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not Path(fc).exists():
        return []
    src = PKG / "fortran"
    out = src / "bin"
    out.mkdir(exist_ok=True)
    built = []
    import tempfile
    for prec, define in (("sp", []), ("dp", ["-DDOUBLE_P"])):
        with tempfile.TemporaryDirectory() as tmp:
            steps = [
                [fc, "-cpp", *define, "-O2", "-module-dir", tmp, "-c", str(src / "bgk3d_gpu_mod.F90"), "-o",
                 f"{tmp}/mod.o"],
                [fc, "-cpp", *define, "-O2", "-module-dir", tmp, f"-I{tmp}", "-c", str(src / "bgk3d_host.F90"), "-o",
                 f"{tmp}/host.o"],
                [fc, f"{tmp}/mod.o", f"{tmp}/host.o", f"-L{LIBDIR}", f"-lbgk3d_gpu_{prec}",
                 "-Wl,-rpath,$ORIGIN/../../lib", "-o", str(out / f"bgk3d_host_{prec}.x")],
            ]
            for cmd in steps:
                if verbose:
                    print(" ".join(cmd), flush=True)
                r = subprocess.run(cmd, capture_output=True, text=True)
                if r.returncode != 0:
                    raise RuntimeError(f"Fortran host build failed:\n{r.stderr}")
        built.append(out / f"bgk3d_host_{prec}.x")
    return built


def build_all(verbose=False):
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=len(TARGETS)) as ex:
        futs = [ex.submit(build_one, n, s, e, verbose) for n, (s, e) in TARGETS.items()]
        libs = [f.result() for f in futs]
    return libs + build_fortran(verbose)


if __name__ == "__main__":
    for p in build_all(verbose=True):
        print("built", p)
