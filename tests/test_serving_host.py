"""Host-side checks of the serving layer (no GPU needed)."""
import numpy as np
import pytest

from bgk3d_gpu_amd import runtime as rt
from bgk3d_gpu_amd import serving as S


def layout_bytes(n, esize):
    # rows padded to 128-byte lines, start shifted so i = 1 opens a line,
    # 19 population planes per z level in each of the two lattices, request
    # region rounded to 64 KiB
    A = 128 // esize
    nx = -(-(n + 2) // A) * A
    plane = nx * (n + 2)
    lstride = -(-(A + 19 * plane * (n + 2)) // A) * A
    b = lstride * 2 * esize
    return -(-b // 65536) * 65536


@pytest.mark.parametrize("n,storage,esize", [(32, "f32", 4), (33, "f16", 2), (130, "f64", 8), (192, "bf16", 2)])
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
