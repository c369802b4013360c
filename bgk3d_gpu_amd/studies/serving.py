"""Continuous-batching serving study: lid-driven cavity requests.

Workload: `count` requests, n ~ U{32..192}, Re ~ logU[100, 3200] (lid
velocity 0.1, nu = 0.1 n / Re), FP32 or FP16 with stochastic rounding with
probability 1/2, 5000 steps each; sizes from default_rng(0), arrivals from
default_rng(1).

1. Offline (all requests at t = 0): the plain sequential loop (one engine
   handle per request, run to completion) and the batching server under an
   HBM budget; throughput in cell updates per second (MLUPS) and the ratio.
2. Online: Poisson arrivals at `load` times the sequential loop's capacity;
   p50 / p99 latency (completion - arrival) of both schedulers.
Requests whose solution became non-finite are listed as unstable.

    python -m bgk3d_gpu_amd.studies.serving [--count 32] [--steps 5000] [--budget-gb 8] > out.json
"""
from __future__ import annotations

import argparse
import json
import sys

from .. import serving as S


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def strip(res):
    return {k: v for k, v in res.items() if k != "records"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=32)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--budget-gb", type=float, default=8.0)
    ap.add_argument("--chunk", type=int, default=100)
    ap.add_argument("--load", type=float, default=0.8)
    ap.add_argument("--no-online", action="store_true")
    a = ap.parse_args(argv)

    budget = int(a.budget_gb * (1 << 30))
    reqs = S.make_workload(a.count, a.steps)
    work = sum(r.cells * r.steps for r in reqs)
    log(f"{a.count} requests, {work / 1e9:.1f} G cell updates, "
        f"{sum(r.storage == 'f16' for r in reqs)} FP16; budget {a.budget_gb} GiB")
    # warm up kernels and the allocator
    S.run_batched(S.make_workload(2, 10, size_seed=7), budget, a.chunk)
    S.run_sequential(S.make_workload(2, 10, size_seed=7))

    seq = S.run_sequential(reqs, progress=log)
    log(f"sequential offline: {seq['throughput_mlups']:.0f} MLUPS, makespan {seq['makespan_s']:.2f} s")
    bat = S.run_batched(reqs, budget, a.chunk, progress=log)
    log(f"batched offline: {bat['throughput_mlups']:.0f} MLUPS, makespan {bat['makespan_s']:.2f} s")
    out = {
        "workload": {"count": a.count, "steps": a.steps, "n_range": [32, 192], "re_range": [100, 3200],
                     "u_lid": 0.1, "p_fp16": 0.5, "size_seed": 0, "arrival_seed": 1,
                     "cell_steps": work,
                     "requests": [dict(n=r.n, re=round(r.re, 1), tau=round(r.tau, 4), storage=r.storage)
                                  for r in reqs]},
        "offline": {"sequential": strip(seq), "batched": strip(bat),
                    "speedup": bat["throughput_mlups"] / seq["throughput_mlups"]},
    }
    # the same request gives the same diagnostics under both schedulers
    same = all((sr["diag"] == br["diag"]).all() or (sr["unstable"] and br["unstable"])
               for sr, br in zip(seq["records"], bat["records"]))
    out["offline"]["diagnostics_identical"] = bool(same)
    if not a.no_online:
        rate = a.load * a.count / seq["makespan_s"]
        reqs_on = S.make_workload(a.count, a.steps, rate=rate)
        log(f"online: Poisson rate {rate:.3f}/s (load {a.load} of the sequential capacity)")
        seq_on = S.run_sequential(reqs_on, online=True, progress=log)
        bat_on = S.run_batched(reqs_on, budget, a.chunk, online=True, progress=log)
        out["online"] = {"rate_per_s": rate, "load": a.load, "sequential": strip(seq_on),
                         "batched": strip(bat_on)}
        log(f"online p50/p99: sequential {seq_on['latency_p50_s']:.2f}/{seq_on['latency_p99_s']:.2f} s, "
            f"batched {bat_on['latency_p50_s']:.2f}/{bat_on['latency_p99_s']:.2f} s")
    print(json.dumps(out, indent=1, default=float))


if __name__ == "__main__":
    main()
