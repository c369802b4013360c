"""Continuous-batching server on the GPU.

* Identity: every request of a mixed batch (sizes that give one, two and
  three waves per row, FP32 / FP64 / FP16 / BF16, nearest and stochastic
  rounding, different step counts so requests join and leave mid-run) ends
  with populations bit-identical to the same request run alone on an engine
  handle, and to the same batch run with another chunk length; also when
  a chunk runs in several sub-batches (more than 3 GB of lattices).
* Admission: with a budget that holds two of the largest requests, no more
  than the budget is ever in use, requests start in arrival order and all finish.
* The diagnostics the server reports equal bgk_diagno of the alone run.
* Reference parity inside a mix: the golden cavity runs of the reference's
  own CPU build (24^3 x 300 and 64^3 x 2000 steps, FP32 and FP64), served
  among FP16/BF16 requests of other sizes, end on the reference's
  populations bit for bit (sha256 of the interior) and report the
  reference's last diagno.dat line (at print precision).
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_support as O
from bgk3d_gpu_amd import runtime as rt
from bgk3d_gpu_amd import serving as S
from bgk3d_gpu_amd.engine import Engine

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


MIX = [  # n, Re, storage, rounding, steps
    (16, 100.0, "f32", 0, 120),
    (40, 400.0, "f16", 1, 200),
    (70, 1000.0, "f32", 0, 90),
    (33, 250.0, "f64", 0, 150),
    (130, 2000.0, "f16", 1, 40),
    (24, 800.0, "bf16", 1, 170),
    (47, 3200.0, "f16", 0, 130),
    (65, 150.0, "bf16", 0, 60),
]


def mix_requests():
    return [S.Request(n, re, st, steps, seed=17 + q, rounding=rnd, keep=True)
            for q, (n, re, st, rnd, steps) in enumerate(MIX)]


def serve(reqs, chunk, budget=8 << 30):
    with S.Server(budget, chunk) as srv:
        ids = [srv.submit(r) for r in reqs]
        done = set()
        while len(done) < len(ids):
            done.update(srv.advance())
        pops = [srv.populations(i, r.n) for i, r in zip(ids, reqs)]
        res = [srv.result(i) for i in ids]
    return pops, res


def alone(r):
    with Engine(r.case(), r.storage, rounding=r.rounding if r.storage in ("f16", "bf16") else 0,
                seed=r.seed) as eng:
        eng.step(r.steps)
        return eng.populations(), eng.diagno()


def test_batched_requests_match_alone_runs(gpu):
    reqs = mix_requests()
    banner(f"serving identity: {len(reqs)} requests, chunk 64")
    pops, res = serve(reqs, 64)
    for q, r in enumerate(reqs):
        banner(f"alone run {q}: n={r.n} {r.storage} rounding={r.rounding} steps={r.steps}")
        p_alone, d_alone = alone(r)
        assert np.array_equal(pops[q], p_alone, equal_nan=True), f"request {q} differs from its alone run"
        assert np.array_equal(res[q]["diag"], d_alone, equal_nan=True)
        assert not res[q]["unstable"]
        assert np.all(np.isfinite(p_alone))
        assert np.any(p_alone != 0.0)     # the lid moved the fluid


def test_requests_across_sub_batches_match_alone_runs(gpu):
    """Requests whose lattices add up to more than one 3 GB sub-batch of a
    chunk (five FP32 cavities of 150^3 and 160^3, 3.4 GB) and two FP16
    ones: the sub-batch split does not change any request's populations."""
    spec = [(160, 900.0, "f32", 0, 23), (150, 300.0, "f32", 0, 31), (160, 2000.0, "f32", 0, 17),
            (150, 700.0, "f32", 0, 29), (160, 1200.0, "f32", 0, 23), (140, 500.0, "f16", 1, 19),
            (150, 800.0, "f16", 1, 27)]
    reqs = [S.Request(n, re, st, steps, seed=40 + q, rounding=rnd, keep=True)
            for q, (n, re, st, rnd, steps) in enumerate(spec)]
    total = sum(S.request_bytes(r) for r in reqs if r.storage == "f32")
    assert total > 3 << 30
    banner(f"serving across sub-batches: {len(reqs)} requests, FP32 lattices {total / 2**30:.2f} GiB")
    pops, res = serve(reqs, 10, budget=16 << 30)
    for q, r in enumerate(reqs):
        banner(f"alone run {q}: n={r.n} {r.storage} steps={r.steps}")
        p_alone, d_alone = alone(r)
        assert np.array_equal(pops[q], p_alone, equal_nan=True), f"request {q} differs from its alone run"
        assert np.array_equal(res[q]["diag"], d_alone, equal_nan=True)


def test_chunk_length_does_not_change_results(gpu):
    reqs = mix_requests()[:5]
    banner("serving chunk 7 vs chunk 1000")
    a, _ = serve(reqs, 7)
    b, _ = serve(reqs, 1000)
    for q in range(len(reqs)):
        assert np.array_equal(a[q], b[q], equal_nan=True)


def test_admission_respects_the_budget(gpu):
    reqs = [S.Request(n, 500.0, st, 60, seed=q) for q, (n, st) in
            enumerate([(96, "f32"), (64, "f16"), (96, "f32"), (80, "f32"), (48, "f16")])]
    need = [S.request_bytes(r) for r in reqs]
    budget = max(need) * 2 + (1 << 16)
    banner(f"serving admission: budget {budget} B, requests {need}")
    with S.Server(budget, 25) as srv:
        ids = [srv.submit(r) for r in reqs]
        done = []
        peak = 0
        while len(done) < len(ids):
            st = srv.stats()
            assert st["bytes_used"] <= st["budget"] == budget
            peak = max(peak, st["active"])
            done += srv.advance()
            peak = max(peak, srv.stats()["active"])
        res = [srv.result(i) for i in ids]
        assert srv.stats()["bytes_used"] == 0
    admits = [r["admit"] for r in res]
    assert admits == sorted(admits), "requests must start in arrival order"
    # the budget holds two of the largest requests: more than one runs at a
    # time, never all five (their sum exceeds the budget)
    assert sum(need) > budget
    assert 2 <= peak < len(reqs)
    for r in res:
        assert r["done"] >= r["admit"] >= r["submit"]
        assert not r["unstable"]


def test_request_larger_than_budget_is_refused(gpu):
    r = S.Request(128, 500.0, "f32", 10)
    with S.Server(S.request_bytes(r) // 2, 10) as srv:
        with pytest.raises(rt.BGKError):
            srv.submit(r)
        small = S.Request(16, 500.0, "f32", 10)
        rid = srv.submit(small)
        while srv.status(rid)[0] != "done":
            srv.advance()
        assert srv.status(rid) == ("done", 10)


def test_shortest_first_policy(gpu):
    """With min_cells = 1 one request runs at a time, least work first; the
    results stay those of the alone runs."""
    reqs = [S.Request(n, 300.0, st, steps, seed=q, keep=True) for q, (n, st, steps) in
            enumerate([(48, "f32", 200), (16, "f16", 100), (32, "f32", 50), (24, "f32", 300)])]
    banner("serving shortest-first policy")
    with S.Server(1 << 30, 1000, policy="shortest", min_cells=1) as srv:
        ids = [srv.submit(r) for r in reqs]
        order = []
        while len(order) < len(ids):
            order += srv.advance()
        pops = [srv.populations(i, r.n) for i, r in zip(ids, reqs)]
    work = [r.cells * r.steps for r in reqs]
    assert order == [ids[q] for q in np.argsort(work, kind="stable")]
    for q, r in enumerate(reqs):
        assert np.array_equal(pops[q], alone(r)[0], equal_nan=True)


FIX = json.loads((Path(__file__).parent / "golden" / "fixtures.json").read_text())


def test_served_requests_reproduce_reference_runs(gpu):
    """SRC/bcond.driven.F90 + SRC/col_MC.F90 runs of the reference's CPU
    build (tests/golden/fixtures.json), served in one mix with 16-bit
    requests of other sizes, chunk 128 (so requests join and leave between
    chunks that cut across the fixtures' step counts)."""
    keys = ["ldc_24_sp", "ldc_24_dp", "ldc_64_re100_sp", "ldc_64_re100_dp"]
    reqs = []
    for q, key in enumerate(keys):
        spec = FIX["runs"][key]
        assert spec["case"] == "ldc" and spec["l"] == spec["m"] == spec["n"]
        reqs.append(S.Request(spec["l"], 0.0, "f32" if spec["prec"] == "sp" else "f64", spec["steps"], seed=q,
                              u0=spec["u0"], keep=True, nu=spec["svisc"]))
    fillers = [S.Request(40, 400.0, "f16", 700, seed=90, rounding=1, keep=True),
               S.Request(97, 1500.0, "bf16", 250, seed=91, rounding=1, keep=True),
               S.Request(33, 200.0, "f16", 1100, seed=92, rounding=0, keep=True)]
    banner(f"serving reference runs {keys} among {len(fillers)} 16-bit requests")
    pops, res = serve(reqs + fillers, 128)
    for q, key in enumerate(keys):
        spec = FIX["runs"][key]
        dt = np.float32 if spec["prec"] == "sp" else np.float64
        P = pops[q].astype(dt)
        assert np.array_equal(P.astype(np.float64), pops[q]), "storage values are exact in float64"
        got = hashlib.sha256(np.ascontiguousarray(P).tobytes()).hexdigest()
        assert got == spec["sha256_interior"], f"{key}: served populations differ from the reference's"
        last = [x for x in spec["texts"]["diagno.dat"].splitlines() if x.strip()][-1]
        O.assert_diagno_close([(spec["steps"], res[q]["diag"])], last)
        assert not res[q]["unstable"]
    for q in range(len(keys), len(reqs) + len(fillers)):
        assert not res[q]["unstable"]


def test_unstable_request_is_retired_early(gpu):
    """A cavity at tau = 0.5 + 3e-5 with lid speed 0.1 blows up within a few
    hundred steps: the step kernel flags its non-finite cells and the server
    retires it as unstable at the end of that chunk, long before its 20,000
    steps; the stable requests it shared chunks with end exactly as their
    alone runs, FP32 and FP16 groups running on the two streams."""
    bad = S.Request(48, 0.0, "f32", 20000, seed=5, keep=False, nu=1e-5)
    good = [S.Request(40, 400.0, "f16", 600, seed=6, rounding=1, keep=True),
            S.Request(56, 300.0, "f32", 500, seed=7, keep=True)]
    banner("serving: retire a non-finite request early")
    with S.Server(4 << 30, 50) as srv:
        bid = srv.submit(bad)
        gids = [srv.submit(r) for r in good]
        done = []
        while len(done) < 3:
            done += srv.advance()
        st, steps = srv.status(bid)
        res = srv.result(bid)
        pops = [srv.populations(i, r.n) for i, r in zip(gids, good)]
        gres = [srv.result(i) for i in gids]
    assert st == "done" and res["unstable"]
    assert 0 < steps < 5000, steps
    assert done.index(bid) < len(done) - 1 or steps < 600
    for q, r in enumerate(good):
        assert not gres[q]["unstable"]
        assert np.array_equal(pops[q], alone(r)[0])


def test_chunk_time_target_keeps_results(gpu):
    """Chunks cut to a time target (50 us: after the first chunk one or a
    few steps each) leave every request's populations bit-identical to its
    alone run."""
    reqs = mix_requests()[:6]
    banner("serving with a 50 us chunk-time target")
    with S.Server(8 << 30, 1000, chunk_ms=0.05) as srv:
        ids = [srv.submit(r) for r in reqs]
        done = set()
        while len(done) < len(ids):
            done.update(srv.advance())
        chunks = srv.stats()["chunks"]
        pops = [srv.populations(i, r.n) for i, r in zip(ids, reqs)]
    assert chunks > 20
    for q, r in enumerate(reqs):
        assert np.array_equal(pops[q], alone(r)[0], equal_nan=True), q
