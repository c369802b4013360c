#!/usr/bin/env python3
"""Benchmark of the D3Q19 BGK time step on one MI355X.

Workload (the bandwidth-roofline case): 3-D Taylor-Green vortex, periodic
256^3, FP32 storage and arithmetic, nu = 0.01 (tau = 0.53), U0 = 0.05, with
the matching pressure field in the initial state.  A "step" is one full time
step: the periodic boundary (folded into the pull) + fused pull-stream + BGK
collide, the reference's loop body (SRC/bgk3d.F90:169-196).

value  = whole-job MLUPS = l*m*n*steps / time / 1e6 (SRC/finalize.F90:106-107)
roofline: dominant kernel (the fused collide), algorithmic bytes per launch =
          38 * 4 B * l*m*n (19 loads + 19 stores per interior cell), divided
          by the kernel's average duration measured with hipEvents around each
          launch in the timed region, against the 8 TB/s HBM peak.
cpu_baseline: the reference's own Fortran CPU routines (oracle/_ref): its
          OpenMP loops on the host's cores, and its serial build, on a
          bounded sample of the same workload.
forms:    FP64 / FP16 / BF16 storage, the reference's mixed-precision forms,
          the reference-compatible ABI, and the 1024x384x384 sphere.
serving:  a reduced stream of lid-driven cavity requests (n in 32..192,
          Re in 100..3200, FP32 or FP16) served by the plain sequential loop
          and by the batching server: cell updates per second offline, and
          p50 / p99 latency under Poisson arrivals.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

METRIC = "MLUPS + % HBM roofline (FP32 vs FP16-store); sweep-server p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stream_bandwidth(lib, rt, nbytes=4 << 30, reps=10):
    """Measured copy bandwidth (read + write bytes / time) of a plain copy kernel."""
    n = nbytes // 4
    src, dst = rt.DeviceBuffer(nbytes, lib), rt.DeviceBuffer(nbytes, lib)
    try:
        lib.bgk_stream_copy(src.ptr, dst.ptr, n)
        e0, e1 = rt.Event(lib), rt.Event(lib)
        e0.record()
        for _ in range(reps):
            lib.bgk_stream_copy(src.ptr, dst.ptr, n)
        e1.record()
        ms = e0.elapsed_ms(e1)
        return 2.0 * nbytes * reps / (ms * 1e-3) / 1e9
    finally:
        src.close()
        dst.close()


def host_threads():
    """Cores this process may run on (the box's CPU share), at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def run_reference(build, case, steps, threads):
    """The reference's routines (oracle/_ref/<build>) timed by their own loop
    timer (SRC/finalize.F90:106-107): MLUPS, or None if the build is absent."""
    exe = REPO / "oracle" / "_ref" / build / "ref_harness.x"
    if not exe.exists():
        return None
    with tempfile.TemporaryDirectory() as d:
        out = subprocess.run([str(exe), d, str(case.l), str(case.m), str(case.n), repr(case.svisc),
                              repr(case.u0), str(steps), str(steps + 1), "1", "0", "0", "3"],
                             capture_output=True, text=True, timeout=600,
                             env=dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close"))
    for line in out.stdout.splitlines():
        if line.startswith("# Mlups"):
            return float(line.split()[-1])
    raise RuntimeError(f"reference harness {build} failed: {out.stdout[-500:]} {out.stderr[-500:]}")


def cpu_baseline(case, steps):
    """The reference's multicore CPU build on the host's cores, beside its
    serial build (bounded sample of the same workload).  The multicore form
    is the reference's OpenMP loops (-DOFFLOAD, private temporaries) run on
    the host's threads: its `do concurrent` build mapped to OpenMP shares the
    loop temporaries between threads and races (oracle/Makefile)."""
    threads = host_threads()
    t0 = time.time()
    omp = run_reference("periodic_sp_omphost", case, steps, threads)
    serial = run_reference("periodic_sp", case, steps, 1)
    if omp is not None and serial is not None:
        return {"value": omp, "unit": "MLUPS", "cores": threads, "kind": "reference",
                "serial_value": serial, "serial_cores": 1, "host_cpus": os.cpu_count(),
                "sample": f"reference Fortran CPU routines, Taylor-Green {case.l}x{case.m}x{case.n} FP32, "
                          f"{steps} steps: OpenMP loops of its offload build run on {threads} host threads "
                          f"(amdflang -O3 -fopenmp -DOFFLOAD, no offload target) and its serial build; "
                          f"wall {time.time() - t0:.1f}s incl. init"}
    # fall back to the C oracle with host threads
    sys.path.insert(0, str(REPO / "tests"))
    import oracle_support as O
    threads = min(16, os.cpu_count() or 1)
    A = O.new_lattice(case.l, case.m, case.n, "sp")
    B = O.new_lattice(case.l, case.m, case.n, "sp")
    O.init(A, case.l, case.m, case.n, "sp", mode=1, u00=case.u0)
    t0 = time.perf_counter()
    O.run("periodic", A, B, case.l, case.m, case.n, steps, prec="sp", svisc=case.svisc, u0=case.u0,
          nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": case.cells * steps / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"C oracle, OpenMP {threads} threads, Taylor-Green {case.l}^3 FP32, {steps} steps"}


def load_traffic(workload, kernel):
    """HBM bytes per launch of the collide from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (profiles/traffic.json, keyed by workload
    and kernel, made by the command in the kernel's profile under
    profiles/); not measured in this run.  None when no pass of this kernel
    on this workload is recorded."""
    p = REPO / "profiles" / "traffic.json"
    if p.exists():
        data = json.loads(p.read_text())
        return data.get(f"{workload}:{kernel}")
    return None


def bench_v1(case, prec, steps, warmup, whole_lines=False):
    """Reference-compatible ABI driven the reference's way (one
    bcond_periodic_gpu + col_mc_gpu per step, caller-owned arrays); total
    milliseconds of the timed steps.  whole_lines: with BGK3D_WHOLE_LINES=1,
    which lets col_mc_gpu store zeros into the halo and padding cells between
    interior rows so that it writes whole 128-byte lines."""
    from bgk3d_gpu_amd.studies.reference_abi import time_steps

    if not whole_lines:
        return time_steps(case, prec, steps, warmup)
    old = os.environ.get("BGK3D_WHOLE_LINES")
    os.environ["BGK3D_WHOLE_LINES"] = "1"
    try:
        return time_steps(case, prec, steps, warmup)
    finally:
        if old is None:
            del os.environ["BGK3D_WHOLE_LINES"]
        else:
            os.environ["BGK3D_WHOLE_LINES"] = old


def bench_engine(case, storage, steps, warmup, want_ke=False, rounding=None, **kw):
    from bgk3d_gpu_amd.engine import Engine

    if rounding is None:
        rounding = 1 if storage in ("f16", "bf16") and not kw.get("mixed") else 0
    with Engine(case, storage, rounding=rounding, **kw) as eng:
        ke0 = eng.diagno()[4] if want_ke else None
        eng.step(warmup)
        eng.sync()
        total_ms, kern_ms = eng.time_steps(steps)
        diag = eng.diagno()
        info = eng.info()
    if want_ke:
        info["ke0"] = ke0
    return total_ms, kern_ms, diag, info


def bench_sphere(steps):
    """Flow past a sphere, 1024 x 384 x 384 FP32 (Re = 300, R = 32 at
    (256, 192, 192), inflow 0.05, solid mask): the handle API, and the
    reference-compatible ABI driven the reference's way (bcond_inflow_gpu +
    bcond_obs_gpu + col_mc_gpu per step, with the reference's shell mask at
    its own centre and u_inflow, the case its build runs)."""
    from bgk3d_gpu_amd.cases import Case

    l, m, n, R, U = 1024, 384, 384, 32, 0.05
    case = Case("inflow", l, m, n, svisc=U * 2 * R / 300.0, u0=U, obstacle=True)
    t, k, d, info = bench_engine(case, "f32", steps, 3, obstacle=2, centre=(l // 4, m // 2, n // 2), radius=R,
                                 u_inflow=U)
    out = {"sphere_engine": {"mlups": round(case.cells * steps / (t * 1e-3) / 1e6, 1),
                             "collide_us": round(k * 1e3, 1), "device_bytes": info["bytes"],
                             "alg_GBs": round(38 * 4 * case.cells / (k * 1e-3) / 1e9, 1),
                             "workload": "sphere 1024x384x384 FP32, solid R=32, Re=300"}}
    log(f"sphere engine: {out['sphere_engine']}")
    ref_case = Case("inflow", l, m, n, svisc=0.02, u0=0.1, obstacle=True)
    t = bench_v1(ref_case, "sp", steps, 3)
    out["sphere_v1_abi"] = {"mlups": round(ref_case.cells * steps / (t * 1e-3) / 1e6, 1),
                            "ms_per_step": round(t / steps, 3),
                            "workload": "sphere 1024x384x384 FP32, reference shell mask, reference layout"}
    log(f"sphere v1: {out['sphere_v1_abi']}")
    return out


def serving_ceiling(reqs, steps=200):
    """Upper bound on the mix's cell updates per second: every request at the
    rate of one large (192^3) cavity of its storage type alone on the GPU,
    which already streams at the HBM roofline -- the cell-weighted harmonic
    mean of those rates.  No scheduler of these kernels can beat it."""
    from bgk3d_gpu_amd import serving as S
    from bgk3d_gpu_amd.engine import Engine

    rates = {}
    for st in sorted({r.storage for r in reqs}):
        r = S.Request(192, 1000.0, st, steps)
        with Engine(r.case(), st, rounding=1 if st in ("f16", "bf16") else 0) as eng:
            eng.step(10)
            eng.sync()
            tot, _ = eng.time_steps(steps)
        rates[st] = r.cells * steps / (tot * 1e-3) / 1e6
    work = sum(r.cells * r.steps for r in reqs)
    ceiling = work / sum(r.cells * r.steps / rates[r.storage] for r in reqs)
    return ceiling, rates


def bench_serving(count, steps, budget_gb=32.0, load=0.8, min_cells=4 << 20, chunk_ms=20.0):
    """Offline cell updates per second and online p50/p99 latency of the
    sequential plain loop -- the reference-compatible ABI driven the
    reference's way, one request after the other -- of a sequential loop of
    engine handles, and of the batching server (shortest remaining work
    first with aging, at least `min_cells` cells per step, chunks cut to
    about `chunk_ms`, non-finite requests retired when detected) on a
    reduced request mix."""
    from bgk3d_gpu_amd import serving as S

    budget = int(budget_gb * (1 << 30))
    reqs = S.make_workload(count, steps)
    S.run_batched(S.make_workload(2, 10, size_seed=7), budget)  # warm-up
    S.run_sequential(S.make_workload(2, 10, size_seed=7), path="v1")
    ceiling, rates = serving_ceiling(reqs)
    plain = S.run_sequential(reqs, path="v1")
    seq = S.run_sequential(reqs)
    kw = dict(policy="shortest", min_cells=min_cells, chunk_ms=chunk_ms)
    bat = S.run_batched(reqs, budget, **kw)
    rate = load * count / plain["makespan_s"]
    on = S.make_workload(count, steps, rate=rate)
    res_on = {"plain": S.run_sequential(on, online=True, path="v1"), "engine_loop": S.run_sequential(on, online=True),
              "batched": S.run_batched(on, budget, online=True, **kw)}
    r3 = lambda x: round(float(x), 3)  # noqa: E731
    evicted = [(q, rec["steps_run"]) for q, rec in enumerate(bat["records"]) if rec["unstable"]]
    return {
        "requests": count, "steps": steps, "cell_steps": int(seq["cell_steps"]),
        "mlups": {"plain": round(plain["throughput_mlups"], 1), "engine_loop": round(seq["throughput_mlups"], 1),
                  "batched": round(bat["throughput_mlups"], 1)},
        "executed_mlups": {"plain": round(plain["executed_mlups"], 1),
                           "engine_loop": round(seq["executed_mlups"], 1),
                           "batched": round(bat["executed_mlups"], 1)},
        "throughput_ratio_vs_plain": r3(bat["throughput_mlups"] / plain["throughput_mlups"]),
        "throughput_ratio_vs_engine_loop": r3(bat["throughput_mlups"] / seq["throughput_mlups"]),
        "ceiling_mlups": round(ceiling, 1),
        "ceiling_note": "every request at the rate of a 192^3 cavity of its storage alone (roofline-bound): "
                        + ", ".join(f"{k} {v:.0f}" for k, v in rates.items())
                        + " MLUPS; the 3x-the-plain-loop target would need "
                        + f"{3 * plain['throughput_mlups']:.0f}",
        "ceiling_over_plain": r3(ceiling / plain["throughput_mlups"]),
        "batched_over_ceiling": r3(bat["executed_mlups"] / ceiling),
        "unstable_requests": len(seq["unstable"]),
        "unstable_retired_at_steps": [s_ for _, s_ in evicted],
        "offline_p50_s": {"plain": r3(plain["latency_p50_s"]), "engine_loop": r3(seq["latency_p50_s"]),
                          "batched": r3(bat["latency_p50_s"])},
        "online_rate_per_s": r3(rate),
        "p50_s": {k: r3(v["latency_p50_s"]) for k, v in res_on.items()},
        "p99_s": {k: r3(v["latency_p99_s"]) for k, v in res_on.items()},
        "plain": "reference-compatible ABI, the reference's loop (bcond + col_mc per step), FP16 requests in FP32",
        "policy": f"shortest remaining work first with aging, >= {min_cells} cells per step, chunks of about "
                  f"{chunk_ms} ms",
        "throughput_definition": "requested cell updates / makespan (a request retired as unstable counts as "
                                 "served); executed_mlups counts the updates actually run",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--prec", default="sp", choices=["sp", "dp"])
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-stream", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the FP64/FP16/BF16/v1 side measurements")
    ap.add_argument("--no-sphere", action="store_true", help="skip the 1024x384x384 sphere forms")
    ap.add_argument("--serving-count", type=int, default=16, help="requests in the serving leg (0: skip)")
    ap.add_argument("--serving-steps", type=int, default=5000)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    from bgk3d_gpu_amd import runtime as rt
    from bgk3d_gpu_amd.cases import Case

    n = args.size
    case = Case("periodic", n, n, n, svisc=0.01, u0=0.05, init=2)  # Taylor-Green with its pressure field
    storage = "f32" if args.prec == "sp" else "f64"
    workload = f"taylor_green_{n}^3_{'fp32' if args.prec == 'sp' else 'fp64'}"
    log(f"workload {workload} steps={args.steps} warmup={args.warmup}")
    total_ms, kern_avg_ms, diag, info = bench_engine(case, storage, args.steps, args.warmup, want_ke=True)
    if not np.all(np.isfinite(diag)):
        raise SystemExit(f"non-finite diagnostics {diag}")
    # kinetic energy against the viscous decay exp(-2 nu |k|^2 t), |k|^2 = 3 (2 pi / n)^2
    t_end = args.warmup + args.steps
    ke_ratio = diag[4] / info["ke0"]
    ke_analytic = float(np.exp(-2 * case.svisc * 3 * (2 * np.pi / n) ** 2 * t_end))
    s_bytes = 4 if args.prec == "sp" else 8
    ms_per_step = total_ms / args.steps
    mlups = case.cells * args.steps / (total_ms * 1e-3) / 1e6
    alg_bytes = 38 * s_bytes * case.cells
    achieved = alg_bytes / (kern_avg_ms * 1e-3) / 1e9
    log(f"MLUPS {mlups:.0f}; collide avg {kern_avg_ms * 1e3:.1f} us -> {achieved:.0f} GB/s")
    forms = {}
    if not args.no_forms:
        fsteps = max(10, args.steps // 4)
        # 16-bit forms store with stochastic rounding (the accurate form);
        # *_rn is nearest-even rounding
        # mixed1 / mixed2: the reference's MIXED1 (FP16 storage, reference-order
        # FP32 collide, nearest) and MIXED2 (FP32 storage, FP64 collide) builds
        for name, st, rnd, sb, mixed in (("f64", "f64", 0, 8, 0), ("f16", "f16", 1, 2, 0),
                                         ("f16_rn", "f16", 0, 2, 0), ("bf16", "bf16", 1, 2, 0),
                                         ("bf16_rn", "bf16", 0, 2, 0), ("f16_mixed1", "f16", 0, 2, 1),
                                         ("f32_mixed2", "f32", 0, 4, 2)):
            if st == storage and not mixed:
                continue
            t, k, d, _ = bench_engine(case, st, fsteps, 5, rounding=rnd, mixed=mixed)
            forms[name] = {"mlups": round(case.cells * fsteps / (t * 1e-3) / 1e6, 1),
                           "collide_us": round(k * 1e3, 2),
                           "alg_GBs": round(38 * sb * case.cells / (k * 1e-3) / 1e9, 1)}
            log(f"{name}: {forms[name]}")
        t = bench_v1(case, args.prec, fsteps, 5)
        forms["v1_abi_" + storage] = {"mlups": round(case.cells * fsteps / (t * 1e-3) / 1e6, 1),
                                      "ms_per_step": round(t / fsteps, 4),
                                      "note": "reference layout (nx=l+2), halo kernels per step"}
        log(f"v1: {forms['v1_abi_' + storage]}")
        t = bench_v1(case, args.prec, fsteps, 5, whole_lines=True)
        forms["v1_abi_" + storage + "_whole_lines"] = {
            "mlups": round(case.cells * fsteps / (t * 1e-3) / 1e6, 1), "ms_per_step": round(t / fsteps, 4),
            "note": "as v1_abi with BGK3D_WHOLE_LINES=1: col_mc_gpu also stores zeros into the halo and padding "
                    "cells between interior rows (values no routine reads), so it writes whole lines"}
        log(f"v1 whole lines: {forms['v1_abi_' + storage + '_whole_lines']}")
        if "f16" in forms:
            forms["f16_over_" + storage] = round(forms["f16"]["mlups"] / mlups, 3)
        if not args.no_sphere:
            forms.update(bench_sphere(max(5, fsteps // 2)))
    stream = None
    if not args.no_stream:
        stream = stream_bandwidth(rt.v1("sp"), rt)
        log(f"measured copy bandwidth {stream:.0f} GB/s")
    serving = None
    if args.serving_count > 0:
        log(f"serving: {args.serving_count} requests x {args.serving_steps} steps ...")
        serving = bench_serving(args.serving_count, args.serving_steps)
        log(f"serving: {serving}")
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(case, args.cpu_steps)
        log(f"cpu baseline {cpu['value']:.2f} MLUPS ({cpu['kind']}, {cpu['cores']} cores)")
    compact = info["nx"] == n + 2  # the engine chose compact rows
    kernel = "collide_compact" if compact else "collide_fused"
    traffic = load_traffic(workload, kernel)
    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.prec == "sp" else "f64",
        "data": "synthetic (analytic Taylor-Green initial state, no dataset)",
        "config": {"workload": workload, "grid": [n, n, n], "boundary": "periodic", "tau": 0.53, "U0": 0.05,
                   "path": "handle API: " + ("compact rows (nx = l + 2)" if compact else "rows padded to 128-byte "
                                              "lines") + ", periodic boundary folded into the fused step",
                   "layout_nx": info["nx"], "device_bytes": info["bytes"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": f"profiles/traffic.json: rocprofv3 FETCH_SIZE and WRITE_SIZE passes of this "
                                       f"kernel (profiles/{kernel}_taylor_green_256_fp32.md), not this run",
                     "traffic_over_alg": round(traffic / alg_bytes, 4) if traffic else None,
                     "kernel": kernel, "kernel_avg_us": round(kern_avg_ms * 1e3, 2),
                     "alg_bytes_per_launch": alg_bytes,
                     "measured_copy_GBs": round(stream, 1) if stream else None,
                     "frac_of_measured_copy": round(achieved / stream, 4) if stream else None},
        "cpu_baseline": cpu,
        "forms": forms,
        "ke_decay": {"steps": t_end, "measured": round(float(ke_ratio), 6), "viscous_analytic": round(ke_analytic, 6)},
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
