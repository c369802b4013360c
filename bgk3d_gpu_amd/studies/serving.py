"""Continuous-batching serving study: lid-driven cavity requests.

Workload: `count` requests, n ~ U{32..192}, Re ~ logU[100, 3200] (lid
velocity 0.1, nu = 0.1 n / Re), FP32 or FP16 with stochastic rounding with
probability 1/2, 5000 steps each; sizes from default_rng(0), arrivals from
default_rng(1).

1. Per size: one request run alone on an engine handle (the plain loop's
   unit of work) against many requests of that size in the batching server.
2. Offline (all requests at t = 0): sequential plain solves -- through the
   reference-compatible ABI the reference's way (16-bit requests in FP32,
   the storage that ABI has), and on one engine handle per request -- and
   the server with both policies ('all': every admitted request advances
   each chunk; 'shortest': least remaining work first, at least `min_cells`
   cells per step); throughput in cell updates per second (MLUPS).
3. Online: Poisson arrivals at `load` times the plain form's capacity;
   p50 / p99 latency (completion - arrival) of every scheduler.
Requests whose solution became non-finite are listed as unstable (the same
requests under every scheduler).

    python -m bgk3d_gpu_amd.studies.serving [--count 32] [--steps 5000] [--budget-gb 32] > out.json
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from .. import serving as S
from ..engine import Engine


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def strip(res):
    return {k: v for k, v in res.items() if k != "records"}


def per_size(budget, steps=200, sizes=(32, 64, 96, 128, 192), storages=("f32", "f16")):
    rows = []
    for st in storages:
        for n in sizes:
            r = S.Request(n, 1000.0, st, steps)
            with Engine(r.case(), st, rounding=1 if st == "f16" else 0) as e:
                e.step(20)
                e.sync()
                tot, kern = e.time_steps(steps)
            single = n ** 3 * steps / (tot * 1e-3) / 1e6
            count = int(min(64, max(1, 2e8 // n ** 3)))
            reqs = [S.Request(n, 1000.0, st, steps, seed=q) for q in range(count)]
            res = S.run_batched(reqs, budget, chunk_steps=steps)
            rows.append({"n": n, "storage": st, "single_mlups": single, "single_step_us": tot * 1e3 / steps,
                         "single_kernel_us": kern * 1e3, "batched_requests": count,
                         "batched_mlups": res["throughput_mlups"], "gain": res["throughput_mlups"] / single})
            log(f"per size {st} n={n}: alone {single:.0f} MLUPS, {count} batched {res['throughput_mlups']:.0f} MLUPS")
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=32)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--budget-gb", type=float, default=32.0)
    ap.add_argument("--chunk", type=int, default=100)
    ap.add_argument("--min-cells", type=int, default=4 << 20)
    ap.add_argument("--chunk-ms", type=float, default=20.0, help="chunk-time target (0: chunk steps only)")
    ap.add_argument("--load", type=float, default=0.8)
    ap.add_argument("--no-online", action="store_true")
    ap.add_argument("--no-per-size", action="store_true")
    a = ap.parse_args(argv)

    budget = int(a.budget_gb * (1 << 30))
    reqs = S.make_workload(a.count, a.steps)
    work = sum(r.cells * r.steps for r in reqs)
    log(f"{a.count} requests, {work / 1e9:.1f} G cell updates, "
        f"{sum(r.storage == 'f16' for r in reqs)} FP16; budget {a.budget_gb} GiB")
    # warm up kernels and the allocator
    S.run_batched(S.make_workload(2, 10, size_seed=7), budget, a.chunk)
    S.run_sequential(S.make_workload(2, 10, size_seed=7))

    out = {"workload": {"count": a.count, "steps": a.steps, "n_range": [32, 192], "re_range": [100, 3200],
                        "u_lid": 0.1, "p_fp16": 0.5, "size_seed": 0, "arrival_seed": 1, "cell_steps": work,
                        "budget_bytes": budget, "chunk_steps": a.chunk, "min_cells": a.min_cells,
                        "requests": [dict(n=r.n, re=round(r.re, 1), tau=round(r.tau, 4), storage=r.storage)
                                     for r in reqs]}}
    if not a.no_per_size:
        out["per_size"] = per_size(budget)

    seq_v1 = S.run_sequential(reqs, path="v1")
    log(f"sequential plain form (v1 ABI) offline: {seq_v1['throughput_mlups']:.0f} MLUPS, "
        f"makespan {seq_v1['makespan_s']:.2f} s")
    seq = S.run_sequential(reqs)
    log(f"sequential engine offline: {seq['throughput_mlups']:.0f} MLUPS, makespan {seq['makespan_s']:.2f} s")
    offline = {"sequential_v1": strip(seq_v1), "sequential_engine": strip(seq)}
    same = True
    for policy in ("all", "shortest"):
        res = S.run_batched(reqs, budget, a.chunk, policy=policy, min_cells=a.min_cells, progress=None,
                            chunk_ms=a.chunk_ms)
        log(f"batched-{policy} offline: {res['throughput_mlups']:.0f} MLUPS ({res['executed_mlups']:.0f} executed), "
            f"makespan {res['makespan_s']:.2f} s, p50 {res['latency_p50_s']:.2f} s, chunks {res['chunks']}, "
            f"launches {res['launches']}")
        offline[f"batched_{policy}"] = strip(res)
        offline[f"speedup_{policy}_vs_v1"] = res["throughput_mlups"] / seq_v1["throughput_mlups"]
        offline[f"speedup_{policy}_vs_engine"] = res["throughput_mlups"] / seq["throughput_mlups"]
        # stable requests only: an unstable one is retired by the server when
        # it turns non-finite, the loop runs it to the end
        same &= all(np.array_equal(sr["diag"], br["diag"])
                    for sr, br in zip(seq["records"], res["records"]) if not sr["unstable"])
    offline["diagnostics_identical_engine_vs_batched"] = bool(same)
    out["offline"] = offline
    if not a.no_online:
        rate = a.load * a.count / seq_v1["makespan_s"]
        reqs_on = S.make_workload(a.count, a.steps, rate=rate)
        log(f"online: Poisson rate {rate:.3f}/s (load {a.load} of the plain form's capacity)")
        online = {"rate_per_s": rate, "load": a.load}
        for name, path in (("sequential_v1", "v1"), ("sequential_engine", "engine")):
            res = S.run_sequential(reqs_on, online=True, path=path)
            online[name] = strip(res)
            log(f"online {name} p50/p99 {res['latency_p50_s']:.2f}/{res['latency_p99_s']:.2f} s")
        for policy in ("all", "shortest"):
            res = S.run_batched(reqs_on, budget, a.chunk, online=True, policy=policy, min_cells=a.min_cells,
                                chunk_ms=a.chunk_ms)
            online[f"batched_{policy}"] = strip(res)
            log(f"online batched-{policy} p50/p99 {res['latency_p50_s']:.2f}/{res['latency_p99_s']:.2f} s")
        out["online"] = online
    print(json.dumps(out, indent=1, default=float))


if __name__ == "__main__":
    main()
