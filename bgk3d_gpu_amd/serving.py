"""Serving many lid-driven cavity solves on one GPU.

`Server` binds the continuous-batching server of include/bgk3d_engine.h
(bgks_*): requests are admitted within an HBM budget (arrival order, or
shortest remaining work first with aging) and the scheduled requests
advance with one kernel launch per storage group, step and sub-batch of at
most 3 GB; a request whose solution turns non-finite is retired at the end
of that chunk.  `run_batched` drives it over a request stream;
`run_sequential` is the plain loop the reference's sweep script stands for
(UTIL/run_sweep.sh:55-63, 128-169): one solve after the other, each with its
own handle (allocate, rest start, `steps` time steps, diagnostics, free).

The workload (`make_workload`): cubic cavities n ~ U{32..192}, Reynolds
number ~ logU[100, 3200] with lid velocity 0.1, so nu = 0.1 n / Re
(SRC/outdat.F90:46) and tau = 0.5 + 3 nu in [0.503, 1.076]; storage FP32 or
FP16 (stochastic rounding) with probability 1/2; Poisson arrivals.  Sizes,
Reynolds numbers and storage come from numpy default_rng(size_seed),
arrival gaps from default_rng(arrival_seed).
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import engine as E
from .cases import Case

STATE = {0: "queued", 1: "active", 2: "done"}
POLICY = {"all": 0, "shortest": 1}


class BGKSRequest(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("svisc", ctypes.c_double), ("u0", ctypes.c_double),
                ("storage", ctypes.c_int), ("rounding", ctypes.c_int), ("steps", ctypes.c_longlong),
                ("seed", ctypes.c_uint), ("keep", ctypes.c_int)]


def lib() -> ctypes.CDLL:
    L = E.lib()
    if not getattr(L, "_bgks_typed", False):
        vp, ci, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
        rq = ctypes.POINTER(BGKSRequest)
        L.bgks_create.argtypes = [ci, ctypes.c_size_t, ci, ctypes.POINTER(vp)]
        L.bgks_destroy.argtypes = [vp]
        L.bgks_set_policy.argtypes = [vp, ci, ll]
        L.bgks_request_bytes.argtypes = [rq, ctypes.POINTER(ctypes.c_size_t)]
        L.bgks_submit.argtypes = [vp, rq, ctypes.POINTER(ll)]
        L.bgks_advance.argtypes = [vp, ctypes.POINTER(ll), ci, ctypes.POINTER(ci)]
        L.bgks_status.argtypes = [vp, ll, ctypes.POINTER(ci), ctypes.POINTER(ll)]
        L.bgks_result.argtypes = [vp, ll, vp, vp, ctypes.POINTER(ci)]
        L.bgks_get_population.argtypes = [vp, ll, ci, vp]
        L.bgks_release.argtypes = [vp, ll]
        L.bgks_stats.argtypes = [vp, ctypes.POINTER(ll)]
        L.bgks_set_chunk_time.argtypes = [vp, ctypes.c_double]
        L.bgks_schedule.argtypes = [ctypes.POINTER(ll), ctypes.POINTER(ll), ctypes.POINTER(ci), ci, ll, ci,
                                    ctypes.POINTER(ci), ctypes.POINTER(ci)]
        L.bgks_clock.argtypes = []
        L.bgks_clock.restype = ctypes.c_double
        L._bgks_typed = True
    return L


def clock() -> float:
    """The server's monotonic clock (seconds)."""
    return lib().bgks_clock()


@dataclass
class Request:
    n: int
    re: float
    storage: str = "f32"
    steps: int = 5000
    seed: int = 0
    arrival: float = 0.0      # seconds after the start of the run
    u0: float = 0.1
    rounding: int = 1         # stochastic rounding for 16-bit storage
    keep: bool = False
    tag: dict = field(default_factory=dict)
    nu: float | None = None   # viscosity given directly (namelist svisc) instead of through re

    @property
    def svisc(self) -> float:
        return self.nu if self.nu is not None else self.u0 * self.n / self.re

    @property
    def tau(self) -> float:
        return 0.5 + 3.0 * self.svisc

    @property
    def cells(self) -> int:
        return self.n ** 3

    def case(self) -> Case:
        return Case("ldc", self.n, self.n, self.n, svisc=self.svisc, u0=self.u0)

    def ctype(self) -> BGKSRequest:
        return BGKSRequest(self.n, self.svisc, self.u0, E.STORAGE[self.storage],
                           self.rounding if self.storage in ("f16", "bf16") else 0, self.steps,
                           self.seed, 1 if self.keep else 0)


def request_bytes(r: Request) -> int:
    out = ctypes.c_size_t()
    E._check(lib().bgks_request_bytes(ctypes.byref(r.ctype()), ctypes.byref(out)), "bgks_request_bytes")
    return out.value


def make_workload(count: int, steps: int = 5000, rate: float | None = None, size_seed: int = 0,
                  arrival_seed: int = 1, n_range=(32, 192), re_range=(100.0, 3200.0), p16: float = 0.5):
    """`count` requests; arrivals at Poisson rate `rate` (per second), all at
    t = 0 if rate is None."""
    g = np.random.default_rng(size_seed)
    ns = g.integers(n_range[0], n_range[1] + 1, count)
    res = np.exp(g.uniform(math.log(re_range[0]), math.log(re_range[1]), count))
    half = g.random(count) < p16
    if rate:
        gaps = np.random.default_rng(arrival_seed).exponential(1.0 / rate, count)
        arrivals = np.cumsum(gaps) - gaps[0]
    else:
        arrivals = np.zeros(count)
    return [Request(int(n), float(re), "f16" if h else "f32", steps, seed=1000 + q, arrival=float(a))
            for q, (n, re, h, a) in enumerate(zip(ns, res, half, arrivals))]


def schedule(remaining, cells, skipped, min_cells: int, max_skips: int = 8) -> list[int]:
    """The server's shortest-first selection (bgks_schedule): indices of the
    requests that advance in the next chunk."""
    n = len(remaining)
    ll, ci = ctypes.c_longlong, ctypes.c_int
    sel, ns = (ci * max(n, 1))(), ci()
    E._check(lib().bgks_schedule((ll * max(n, 1))(*remaining), (ll * max(n, 1))(*cells), (ci * max(n, 1))(*skipped),
                                 n, min_cells, max_skips, sel, ctypes.byref(ns)), "bgks_schedule")
    return [sel[q] for q in range(ns.value)]


class Server:
    """policy 'all': every admitted request advances each chunk, arrival-order
    admission; 'shortest': least remaining work first, enough requests to keep
    `min_cells` cells per step in flight, with aging.  chunk_ms > 0 cuts
    chunks to about that many milliseconds."""

    def __init__(self, budget_bytes: int, chunk_steps: int = 100, device: int = 0, policy: str = "all",
                 min_cells: int = 0, chunk_ms: float = 0.0):
        self.L = lib()
        h = ctypes.c_void_p()
        E._check(self.L.bgks_create(device, budget_bytes, chunk_steps, ctypes.byref(h)), "bgks_create")
        self.h = h.value
        if policy != "all" or min_cells:
            E._check(self.L.bgks_set_policy(self.h, POLICY[policy], min_cells), "bgks_set_policy")
        if chunk_ms:
            E._check(self.L.bgks_set_chunk_time(self.h, chunk_ms), "bgks_set_chunk_time")

    def submit(self, r: Request) -> int:
        rid = ctypes.c_longlong()
        E._check(self.L.bgks_submit(self.h, ctypes.byref(r.ctype()), ctypes.byref(rid)), "bgks_submit")
        return rid.value

    def advance(self) -> list[int]:
        buf = (ctypes.c_longlong * 4096)()
        nd = ctypes.c_int()
        E._check(self.L.bgks_advance(self.h, buf, 4096, ctypes.byref(nd)), "bgks_advance")
        return [buf[q] for q in range(nd.value)]

    def status(self, rid: int):
        st, steps = ctypes.c_int(), ctypes.c_longlong()
        E._check(self.L.bgks_status(self.h, rid, ctypes.byref(st), ctypes.byref(steps)), "bgks_status")
        return STATE[st.value], steps.value

    def result(self, rid: int) -> dict:
        diag, times, unstable = np.zeros(5), np.zeros(3), ctypes.c_int()
        E._check(self.L.bgks_result(self.h, rid, diag.ctypes.data, times.ctypes.data, ctypes.byref(unstable)),
                 "bgks_result")
        return dict(diag=diag, submit=times[0], admit=times[1], done=times[2], unstable=bool(unstable.value))

    def population(self, rid: int, d: int, n: int) -> np.ndarray:
        out = np.empty((n, n, n))
        E._check(self.L.bgks_get_population(self.h, rid, d, out.ctypes.data), "bgks_get_population")
        return out

    def populations(self, rid: int, n: int) -> np.ndarray:
        return np.stack([self.population(rid, d, n) for d in range(19)])

    def release(self, rid: int):
        E._check(self.L.bgks_release(self.h, rid), "bgks_release")

    def stats(self) -> dict:
        out = (ctypes.c_longlong * 8)()
        E._check(self.L.bgks_stats(self.h, out), "bgks_stats")
        keys = ("queued", "active", "done", "bytes_used", "budget", "launches", "cell_steps", "chunks")
        return dict(zip(keys, list(out)))

    def close(self):
        if self.h:
            self.L.bgks_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _summary(requests, records, t0, t_end, label):
    """throughput_mlups: the requested cell updates of the stream over the
    makespan (a request retired as unstable counts as served);
    executed_mlups: the cell updates actually run over the makespan."""
    lat = np.array([rec["done"] - (t0 + r.arrival) for r, rec in zip(requests, records)])
    work = sum(r.cells * r.steps for r in requests)
    ran = sum(r.cells * rec.get("steps_run", r.steps) for r, rec in zip(requests, records))
    return {
        "scheduler": label,
        "requests": len(requests),
        "cell_steps": work,
        "cell_steps_run": ran,
        "makespan_s": t_end - t0,
        "throughput_mlups": work / (t_end - t0) / 1e6,
        "executed_mlups": ran / (t_end - t0) / 1e6,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_mean_s": float(lat.mean()),
        "latencies_s": [float(x) for x in lat],
        "unstable": [q for q, rec in enumerate(records) if rec["unstable"]],
    }


def _wait_until(t):
    while True:
        d = t - clock()
        if d <= 0:
            return
        time.sleep(min(d, 0.05))


def run_batched(requests, budget_bytes: int, chunk_steps: int = 100, online: bool = False,
                progress=None, policy: str = "all", min_cells: int = 0, chunk_ms: float = 0.0) -> dict:
    """Serve `requests` with the batching server.  Offline: all submitted at
    the start.  Online: each submitted at its arrival time (checked between
    chunks, so a request waits at most one chunk to be queued)."""
    order = sorted(range(len(requests)), key=lambda q: requests[q].arrival)
    records = [None] * len(requests)
    with Server(budget_bytes, chunk_steps, policy=policy, min_cells=min_cells, chunk_ms=chunk_ms) as srv:
        ids = {}
        nxt = 0
        t0 = clock()
        finished = 0
        peak_admitted = 0
        while finished < len(requests):
            now = clock()
            while nxt < len(order) and (not online or t0 + requests[order[nxt]].arrival <= now):
                q = order[nxt]
                ids[srv.submit(requests[q])] = q
                nxt += 1
            st = srv.stats()
            if st["queued"] == 0 and st["active"] == 0:
                if nxt >= len(order):
                    raise RuntimeError("server idle with requests unfinished")
                _wait_until(t0 + requests[order[nxt]].arrival)
                continue
            done = srv.advance()
            peak_admitted = max(peak_admitted, srv.stats()["active"] + len(done))
            for rid in done:
                q = ids[rid]
                records[q] = srv.result(rid)
                records[q]["steps_run"] = srv.status(rid)[1]
                finished += 1
                if not requests[q].keep:
                    srv.release(rid)
            if progress and done:
                progress(f"batched: {finished}/{len(requests)} done")
        t_end = clock()
        st = srv.stats()
    out = _summary(requests, records, t0, t_end, f"batched-{policy}")
    out.update(budget_bytes=budget_bytes, chunk_steps=chunk_steps, launches=st["launches"], chunks=st["chunks"],
               peak_admitted=peak_admitted, policy=policy, min_cells=min_cells, chunk_ms=chunk_ms)
    out["records"] = records
    return out


def _solve_plain_form(r: Request):
    """One request through the reference-compatible ABI the reference's way:
    caller-owned arrays in its layout, host f_eq, per step bcond_driven_gpu
    then col_mc_gpu and the pointer swap.  That ABI has FP32 and FP64 only,
    so a 16-bit request runs in FP32."""
    from .simulation import Simulation

    sim = Simulation(r.case(), "dp" if r.storage == "f64" else "sp", poison=False)
    try:
        sim.step(r.steps)
        return np.asarray(sim.diagno(), dtype=float)
    finally:
        sim.close()


def run_sequential(requests, online: bool = False, progress=None, path: str = "engine") -> dict:
    """The plain loop: one request at a time in arrival order, run to
    completion before the next starts.  path 'engine': each request on its own
    engine handle (the same kernels and storage as the server); 'v1': through
    the reference-compatible ABI (`_solve_plain_form`)."""
    order = sorted(range(len(requests)), key=lambda q: requests[q].arrival)
    records = [None] * len(requests)
    t0 = clock()
    for c, q in enumerate(order):
        r = requests[q]
        if online:
            _wait_until(t0 + r.arrival)
        t_start = clock()
        if path == "v1":
            diag = _solve_plain_form(r)
        else:
            with E.Engine(r.case(), r.storage, rounding=r.rounding if r.storage in ("f16", "bf16") else 0,
                          seed=r.seed) as eng:
                eng.step(r.steps)
                diag = eng.diagno()
        t_done = clock()
        records[q] = dict(diag=diag, submit=t0 + r.arrival, admit=t_start, done=t_done,
                          unstable=not bool(np.all(np.isfinite(diag))))
        if progress and (c + 1) % 4 == 0:
            progress(f"sequential: {c + 1}/{len(requests)} done")
    t_end = clock()
    out = _summary(requests, records, t0, t_end, f"sequential-{path}")
    out["records"] = records
    return out


__all__ = ["Request", "Server", "make_workload", "run_batched", "run_sequential", "request_bytes", "clock",
           "BGKSRequest", "schedule"]
