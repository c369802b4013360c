"""Host-side checks of the serving layer (no GPU needed)."""
import numpy as np
import pytest

from bgk3d_gpu_amd import runtime as rt
from bgk3d_gpu_amd import serving as S


def layout_bytes(n, esize):
    # rows padded to 128-byte lines (start shifted so i = 1 opens a line),
    # unless the last line of a row would leave more than 3 % of a row
    # unused (9 % with 16-bit storage) or the lattice has 2^24 cells or more
    # with 32/64-bit storage: then compact rows of n + 2 cells with each
    # plane padded to whole lines; 19 population planes per z level in each
    # of the two lattices, request region rounded to 64 KiB
    A = 128 // esize
    up = lambda v: -(-v // A) * A  # noqa: E731
    wide = esize >= 4
    if (up(n) - n) / n > (0.03 if wide else 0.09) or (wide and n ** 3 >= 1 << 24):
        plane = up((n + 2) * (n + 2))
    else:
        plane = up(n + 2) * (n + 2)
    lstride = up(A + 19 * plane * (n + 2))
    b = lstride * 2 * esize
    return -(-b // 65536) * 65536


@pytest.mark.parametrize("n,storage,esize", [(32, "f32", 4), (33, "f16", 2), (130, "f64", 8), (192, "bf16", 2),
                                             (129, "f32", 4), (180, "f32", 4), (97, "f16", 2), (188, "f32", 4),
                                             (184, "f32", 4), (250, "f16", 2), (120, "f64", 8), (256, "f32", 4),
                                             (256, "f16", 2), (200, "f16", 2), (184, "bf16", 2)])
def test_request_bytes_follow_the_layout(n, storage, esize):
    assert S.request_bytes(S.Request(n, 1000.0, storage)) == layout_bytes(n, esize)


def test_request_bytes_reject_bad_requests():
    with pytest.raises(rt.BGKError):
        S.request_bytes(S.Request(0, 100.0))
    bad = S.Request(16, -5.0)               # negative viscosity
    with pytest.raises(rt.BGKError):
        S.request_bytes(bad)


def test_workload_is_seeded_and_in_range():
    a = S.make_workload(200, rate=3.0)
    b = S.make_workload(200, rate=3.0)
    assert [(r.n, r.re, r.storage, r.arrival) for r in a] == [(r.n, r.re, r.storage, r.arrival) for r in b]
    ns = np.array([r.n for r in a])
    taus = np.array([r.tau for r in a])
    assert ns.min() >= 32 and ns.max() <= 192
    assert taus.min() >= 0.5 + 3 * 0.1 * 32 / 3200 - 1e-12 and taus.max() <= 0.5 + 3 * 0.1 * 192 / 100 + 1e-12
    arr = np.array([r.arrival for r in a])
    assert arr[0] == 0 and np.all(np.diff(arr) >= 0)
    # mean gap close to 1/rate
    assert abs(np.diff(arr).mean() - 1 / 3.0) < 0.1
    frac16 = np.mean([r.storage == "f16" for r in a])
    assert 0.35 < frac16 < 0.65


def test_server_needs_a_gpu():
    from conftest import gpu_present
    if gpu_present():
        pytest.skip("a GPU is present")
    with pytest.raises(rt.BGKError):
        S.Server(1 << 30)


def test_schedule_is_shortest_remaining_work_first():
    rem = [500, 100, 300, 200]
    cells = [10, 10, 10, 10]
    assert S.schedule(rem, cells, [0] * 4, min_cells=1) == [1]
    assert S.schedule(rem, cells, [0] * 4, min_cells=25) == [1, 3, 2]
    assert S.schedule(rem, cells, [0] * 4, min_cells=10 ** 9) == [1, 3, 2, 0]
    # ties keep index order; at least one request always runs
    assert S.schedule([5, 5, 5], [1, 1, 1], [0, 0, 0], min_cells=0) == [0]


def test_schedule_ages_a_passed_over_request():
    # a request passed over max_skips chunks in a row goes first; the
    # shortest ones follow while the cells in flight stay below min_cells
    assert S.schedule([900, 10, 20], [100, 1, 1], [8, 0, 0], min_cells=1, max_skips=8) == [0]
    assert S.schedule([900, 10, 20], [100, 1, 1], [8, 0, 0], min_cells=101, max_skips=8) == [0, 1]
    assert S.schedule([900, 10, 20], [100, 1, 1], [7, 0, 0], min_cells=1, max_skips=8) == [1]


def test_large_request_not_starved_by_a_steady_small_stream():
    """Host-side simulation of the server's shortest-first loop (the
    selection of bgks_schedule, skipped counts kept as bgks_advance keeps
    them): a new small request arrives every chunk and each chunk runs only
    enough requests for min_cells.  Without aging the large request would
    never be picked; with it, it advances at least every 9 chunks and ends."""
    big_cells, small_cells, steps, chunk = 1000, 10, 100, 10
    reqs = [{"id": "big", "cells": big_cells, "left": 200, "skipped": 0}]
    finished, picks = [], []
    for c in range(400):
        reqs.append({"id": c, "cells": small_cells, "left": steps, "skipped": 0})
        sel = S.schedule([r["cells"] * r["left"] for r in reqs], [r["cells"] for r in reqs],
                         [r["skipped"] for r in reqs], min_cells=small_cells)
        for q, r in enumerate(reqs):
            if q in sel:
                r["skipped"] = 0
                r["left"] -= chunk
            else:
                r["skipped"] += 1
        if 0 in sel and reqs[0]["id"] == "big":
            picks.append(c)
        finished += [r["id"] for r in reqs if r["left"] <= 0]
        reqs = [r for r in reqs if r["left"] > 0]
        if "big" in finished:
            break
    assert "big" in finished
    assert max(b - a for a, b in zip([-1] + picks, picks)) <= 9
