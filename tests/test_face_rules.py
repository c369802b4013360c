"""CPU pins of the one-launch boundary rules against the reference's order.

The reference fills halo faces one after the other and runs its obstacle
bounce-back as one in-place loop (SRC/bcond.periodic.F90:105-252,
SRC/bcond.channel.F90:78-281, SRC/bcond.obs.F90:101-122).  The GPU kernels
(bgk3d_gpu_amd/csrc/v1_kernels.hpp: periodic_faces, channel_faces,
obstacle_bb) do each in one launch, every thread writing final values
computed from values no thread of the launch writes.  These tests restate
each kernel's per-thread rule in Python, check that the rule is race-free
(every written (cell, population) has one writer, and no value a thread
reads is written by another), and that applying it to a seeded random state
gives the oracle's sequential result on the whole arrays, halo included.
The oracle's routines are pinned to the compiled reference by
test_oracle_fixtures.py (single-routine pins on random states).
"""
import numpy as np
import pytest

import oracle_support as O

CX, CY, CZ = O.CX, O.CY, O.CZ
OPP = [11, 12, 9, 10, 13, 14, 15, 16, 17, 2, 3, 0, 1, 4, 5, 6, 7, 8, 18]
ENTER = {("x", -1): [0, 1, 2, 3, 4], ("x", 1): [9, 10, 11, 12, 13],
         ("y", -1): [2, 6, 7, 8, 11], ("y", 1): [0, 9, 15, 16, 17],
         ("z", -1): [3, 5, 6, 12, 17], ("z", 1): [1, 8, 10, 14, 15]}


def side(c, mx):
    return -1 if c == 0 else (1 if c == mx + 1 else 0)


def enters(s, c):
    return (s < 0 and c > 0) or (s > 0 and c < 0)


def wrap(s, mx):
    return mx if s < 0 else 1


def face_rule_pairs(l, m, n, walls, x=True):
    """(target, source) pairs of periodic_faces<x> (walls=False) or
    channel_faces (walls=True); cells as (d, k, j, i).  x=False: the y and z
    faces alone, the x halo copied as it stands (bcond_inflow)."""
    pairs = []
    # x part: rows j = 0..m+1, k = 0..n+1, q = 0..9
    for q in range(10 if x else 0):
        d = ENTER[("x", -1)][q] if q < 5 else ENTER[("x", 1)][q - 5]
        it, is_ = (0, l) if q < 5 else (l + 1, 1)
        for k in range(n + 2):
            for j in range(m + 2):
                sj, sk = side(j, m), side(k, n)
                if enters(sj, CY[d]):
                    continue
                if enters(sk, CZ[d]):
                    if not walls or 1 <= j + CY[d] <= m:
                        continue
                pairs.append(((d, k, j, it), (d, k, j, is_)))
    # y part: (i, k), i = 0..l+1, k = 0..n+1
    for k in range(n + 2):
        for i in range(l + 2):
            si, sk = side(i, l), side(k, n)
            for h, (jt, js) in enumerate(((0, m), (m + 1, 1))):
                for d in ENTER[("y", -1 if h == 0 else 1)]:
                    if enters(sk, CZ[d]):
                        if not walls or 1 <= i + CX[d] <= l:
                            continue
                    is_ = wrap(si, l) if x and enters(si, CX[d]) else i
                    pairs.append(((d, k, jt, i), (d, k, js, is_)))
    if not walls:
        # z part over the full (i, j) face
        for j in range(m + 2):
            for i in range(l + 2):
                si, sj = side(i, l), side(j, m)
                for h, (kt, ks) in enumerate(((0, n), (n + 1, 1))):
                    for d in ENTER[("z", -1 if h == 0 else 1)]:
                        is_ = wrap(si, l) if x and enters(si, CX[d]) else i
                        js = wrap(sj, m) if enters(sj, CY[d]) else j
                        pairs.append(((d, kt, j, i), (d, ks, js, is_)))
    else:
        # walls_z_cell over interior columns: population d entering across
        # the wall plane at (i, j) - c_d, from the cell's opposite population
        for j in range(1, m + 1):
            for i in range(1, l + 1):
                for d in ENTER[("z", -1)]:
                    pairs.append(((d, 0, j - CY[d], i - CX[d]), (OPP[d], 1, j, i)))
                for d in ENTER[("z", 1)]:
                    pairs.append(((d, n + 1, j - CY[d], i - CX[d]), (OPP[d], n, j, i)))
    return pairs


def apply_one_launch(A, pairs):
    targets = [t for t, _ in pairs]
    assert len(set(targets)) == len(targets), "a (cell, population) has two writers"
    tset = set(targets)
    assert not any(s in tset for _, s in pairs), "a thread reads a value another thread writes"
    old = A.copy()
    for t, s in pairs:
        A[t] = old[s]


@pytest.mark.parametrize("case,walls", [("periodic", False), ("channel", True)])
@pytest.mark.parametrize("dims", [(7, 6, 5), (5, 8, 6), (9, 4, 7)])
def test_one_launch_faces_equal_sequential_faces(case, walls, dims):
    l, m, n = dims
    rng = np.random.default_rng(l * 100 + m * 10 + n)
    A = O.new_lattice(l, m, n, "dp", ipad=2, jpad=1)
    A[...] = rng.standard_normal(A.shape)
    ref = A.copy()
    O.bcond(case, ref, l, m, n, "dp")
    ours = A.copy()
    apply_one_launch(ours, face_rule_pairs(l, m, n, walls))
    assert np.array_equal(ours, ref)


@pytest.mark.parametrize("dims", [(7, 6, 5), (5, 8, 6)])
def test_one_launch_yz_faces_equal_sequential_faces(dims):
    """The y and z faces of bcond_inflow after its computed x faces
    (SRC/bcond.inflow.F90:183-328): copied along with the x halo unwrapped."""
    l, m, n = dims
    rng = np.random.default_rng(l + m + n)
    A = O.new_lattice(l, m, n, "dp", ipad=1)
    A[...] = rng.standard_normal(A.shape)
    ref = A.copy()
    O.fn("faces_yz", "dp")(O.ptrs(ref), l, m, n, A.shape[3], A.shape[2])
    ours = A.copy()
    apply_one_launch(ours, face_rule_pairs(l, m, n, False, x=False))
    assert np.array_equal(ours, ref)


def obstacle_rule_pairs(obs, box):
    """(target, source) pairs of obstacle_bb: for solid S of the box and
    direction d, a_d(S) = a_opp(d)(S + c_d) unless S + c_d is a solid cell of
    the box that the reference's loop (k, then j, then i) visits first."""
    imin, imax, jmin, jmax, kmin, kmax = box

    def visited_solid(i, j, k):
        if not (imin <= i <= imax and jmin <= j <= jmax and kmin <= k <= kmax):
            return False
        return obs[k - 1, j - 1, i - 1] == 1

    pairs = []
    for k in range(kmin, kmax + 1):
        for j in range(jmin, jmax + 1):
            for i in range(imin, imax + 1):
                if not visited_solid(i, j, k):
                    continue
                for d in range(18):
                    dx, dy, dz = CX[d], CY[d], CZ[d]
                    earlier = dz < 0 or (dz == 0 and (dy < 0 or (dy == 0 and dx < 0)))
                    if earlier and visited_solid(i + dx, j + dy, k + dz):
                        continue
                    pairs.append(((d, k, j, i), (OPP[d], k + dz, j + dy, i + dx)))
    return pairs


@pytest.mark.parametrize("dims", [(40, 24, 24), (30, 20, 22)])
def test_one_pass_obstacle_rule_equals_serial_loop(dims):
    l, m, n = dims
    obs, box, _ = O.build_obs(l, m, n)
    assert obs.sum() > 100
    rng = np.random.default_rng(7)
    A = O.new_lattice(l, m, n, "dp")
    A[...] = rng.standard_normal(A.shape)
    ref = A.copy()
    f = O.fn("bcond_obs", "dp")
    f(O.ptrs(ref), obs.ctypes.data_as(O.ctypes.c_void_p), l, m, n, A.shape[3], A.shape[2], *box)
    ours = A.copy()
    apply_one_launch(ours, obstacle_rule_pairs(obs, box))
    assert np.array_equal(ours, ref)
    # the rule differs from skipping every solid neighbour (the solid cells'
    # values would then differ from the serial loop's)
    assert any(obs[s[1] - 1, s[2] - 1, s[3] - 1] == 1 for _, s in obstacle_rule_pairs(obs, box)
               if 1 <= s[1] <= n and 1 <= s[2] <= m and 1 <= s[3] <= l)
