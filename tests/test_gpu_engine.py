"""Handle API (library-owned lattices, boundaries folded into the pull) vs
the CPU oracle.

FP32/FP64 storage must be bit-identical to the oracle on every fluid interior
cell (the stated tolerance of 1e-5 / 1e-12 relative is asserted too).  The
16-bit forms are checked against the FP64 engine run of the same case.
"""
import sys

import numpy as np
import pytest

import oracle_support as O
from bgk3d_gpu_amd.cases import Case
from bgk3d_gpu_amd.engine import Engine

pytestmark = pytest.mark.gpu

REL_TOL = {"f32": 1e-5, "f64": 1e-12}
PREC = {"f32": "sp", "f64": "dp"}

CASES = {
    "ldc": Case("ldc", 20, 16, 12, svisc=0.064, u0=0.1),
    "ldc_odd": Case("ldc", 67, 13, 9, svisc=0.02, u0=0.1),
    "periodic": Case("periodic", 20, 16, 12, svisc=0.01, u0=0.05, init=1),
    "periodic_odd": Case("periodic", 70, 11, 7, svisc=0.01, u0=0.05, init=1),
    "channel": Case("channel", 24, 16, 12, svisc=0.1, u0=0.02048),
    "inflow": Case("inflow", 40, 24, 24, svisc=0.02, u0=0.1, obstacle=True),
    "inflow_long": Case("inflow", 130, 24, 16, svisc=0.02, u0=0.1, obstacle=True),
    # forms the reference's fixtures do not cover: obstacles under periodic
    # boundaries, the Smagorinsky relaxation with obstacles and on odd sizes
    "periodic_obstacle": Case("periodic", 40, 24, 24, svisc=0.01, u0=0.05, init=1, obstacle=True),
    "inflow_les": Case("inflow", 40, 24, 24, svisc=0.002, u0=0.1, obstacle=True, les=True),
    "periodic_obstacle_les": Case("periodic", 40, 24, 24, svisc=0.001, u0=0.05, init=1, obstacle=True, les=True),
    "ldc_odd_les": Case("ldc", 67, 13, 9, svisc=0.002, u0=0.1, les=True),
}


def banner(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def oracle_state(case, prec, steps):
    A = O.new_lattice(case.l, case.m, case.n, prec)
    B = O.new_lattice(case.l, case.m, case.n, prec)
    O.init(A, case.l, case.m, case.n, prec, mode=case.init, u00=case.u00)
    obs = box = None
    if case.obstacle:
        obs, box, _ = O.build_obs(case.l, case.m, case.n)
    O.run(case.boundary, A, B, case.l, case.m, case.n, steps, prec=prec, svisc=case.svisc, u0=case.u0, obs=obs,
          box=box, les=case.les, cte_s=case.cte_s)
    return A, obs


@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("storage", ["f32", "f64"])
@pytest.mark.parametrize("name", list(CASES))
def test_engine_bit_exact(gpu, name, storage, layout):
    """Both row layouts: padded to 128-byte lines (collide_fused) and
    compact (collide_compact; with inflow/outflow its zeroed x halos are
    rewritten before every collide)."""
    case = CASES[name]
    steps = 37
    banner(f"engine {name} {storage} layout={layout} {case.l}x{case.m}x{case.n} steps={steps}")
    with Engine(case, storage, layout=layout) as eng:
        eng.step(steps)
        G = eng.populations()
        dg = eng.diagno()
    prec = PREC[storage]
    A, obs = oracle_state(case, prec, steps)
    oi = A[:, 1:-1, 1:-1, 1:-1].astype(np.float64)
    gi = G
    if obs is not None:
        fluid = obs == 0
        gi, oi = gi[:, fluid], oi[:, fluid]
    assert np.isfinite(gi).all()
    assert np.abs(gi - oi).max() <= REL_TOL[storage] * np.abs(oi).max()
    assert np.array_equal(gi, oi), f"max |diff| {np.abs(gi - oi).max()}"
    if obs is None:
        od = O.diagno(A, case.l, case.m, case.n, prec)
        np.testing.assert_allclose(dg, od, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("storage", ["f32", "f64"])
@pytest.mark.parametrize("layout", [1, 2])
def test_engine_draglift(gpu, storage, layout):
    case = CASES["inflow"]
    banner(f"engine draglift {storage} layout={layout}")
    with Engine(case, storage, layout=layout) as eng:
        eng.step(25)
        G = eng.populations()
        dl = eng.draglift()
    prec = PREC[storage]
    obs, box, radius = O.build_obs(case.l, case.m, case.n)
    # solid cells hold different (unread) values in the two implementations;
    # draglift reads only fluid cells, so compare on the oracle's own state
    A, _ = oracle_state(case, prec, 25)
    fluid = obs == 0
    np.testing.assert_array_equal(G[:, fluid], A[:, 1:-1, 1:-1, 1:-1][:, fluid])
    ref = O.draglift(A, obs, case.l, case.m, case.n, radius, prec=prec)
    tol = 1e-4 if storage == "f32" else 1e-10  # oracle sums in the storage kind, engine in FP64
    np.testing.assert_allclose(dl, ref, rtol=tol, atol=tol * np.abs(ref).max())


def test_engine_fields_consistent(gpu):
    case = CASES["periodic"]
    with Engine(case, "f64") as eng:
        eng.step(10)
        rho, ux, uy, uz = eng.fields()
        P = eng.populations()
    rho_p = P.sum(axis=0) + 1.0
    np.testing.assert_allclose(rho, rho_p, rtol=1e-13)
    ux_p = (P[0] + P[1] + P[2] + P[3] + P[4] - P[9] - P[10] - P[11] - P[12] - P[13]) / rho_p
    np.testing.assert_allclose(ux, ux_p, rtol=1e-10, atol=1e-15)


@pytest.mark.parametrize("storage,rounding", [("f16", 0), ("bf16", 0), ("f16", 1), ("bf16", 1)])
def test_low_precision_close_to_f64(gpu, storage, rounding):
    case = Case("ldc", 32, 32, 32, svisc=0.064, u0=0.1)
    steps = 300
    banner(f"engine {storage} rounding={rounding} vs f64")
    with Engine(case, "f64") as ref:
        ref.step(steps)
        _, ux0, uy0, uz0 = ref.fields()
    with Engine(case, storage, rounding=rounding) as eng:
        eng.step(steps)
        _, ux, uy, uz = eng.fields()
    num = np.sqrt(((ux - ux0) ** 2 + (uy - uy0) ** 2 + (uz - uz0) ** 2).sum())
    den = np.sqrt((ux0 ** 2 + uy0 ** 2 + uz0 ** 2).sum())
    rel = num / den
    print(f"[gpu] {storage} rounding={rounding}: velocity L2 rel error {rel:.3e}", flush=True)
    assert np.isfinite(rel)
    assert rel < (2e-2 if storage == "bf16" else 5e-3)


@pytest.mark.parametrize("storage,rounding,bound", [("f16", 1, 1e-2), ("bf16", 1, 3e-2)])
def test_low_precision_les_close_to_f64(gpu, storage, rounding, bound):
    """The 16-bit forms with the Smagorinsky relaxation (lean collide) against
    the FP64 engine, cavity at Re = 640 (tau = 0.515; at Re = 1600 on 32^3
    the flow diverges in every precision, the oracle's FP64 run included)."""
    case = Case("ldc", 32, 32, 32, svisc=0.005, u0=0.1, les=True)
    steps = 300
    banner(f"engine LES {storage} rounding={rounding} vs f64")
    with Engine(case, "f64") as ref:
        ref.step(steps)
        _, ux0, uy0, uz0 = ref.fields()
    with Engine(case, storage, rounding=rounding) as eng:
        eng.step(steps)
        _, ux, uy, uz = eng.fields()
    num = np.sqrt(((ux - ux0) ** 2 + (uy - uy0) ** 2 + (uz - uz0) ** 2).sum())
    den = np.sqrt((ux0 ** 2 + uy0 ** 2 + uz0 ** 2).sum())
    rel = num / den
    print(f"[gpu] LES {storage} rounding={rounding}: velocity L2 rel error {rel:.3e}", flush=True)
    assert np.isfinite(rel)
    assert rel < bound


@pytest.mark.parametrize("storage,tol", [("f64", 1e-13), ("f32", 5e-6)])
def test_channel_walls_in_y_equal_walls_in_z_swapped(gpu, storage, tol):
    """The channel with walls at y = 1, m (wall_axis 2) is the reference's
    channel (walls at z = 1, n) with y and z exchanged: on a square cross
    section the fields agree after the swap, u_y and u_z exchanged, to the
    rounding of the differently ordered sums (the oracle has only the
    reference orientation; this pins the other one)."""
    base = dict(svisc=0.1, u0=0.02048)
    cz = Case("channel", 24, 16, 16, **base)
    cy = Case("channel", 24, 16, 16, wall_axis=2, **base)
    steps = 400
    banner(f"channel walls y vs z {storage}")
    with Engine(cz, storage) as e:
        e.step(steps)
        rz, uxz, uyz, uzz = e.fields()
    with Engine(cy, storage) as e:
        e.step(steps)
        ry, uxy, uyy, uzy = e.fields()
    T = lambda a: np.swapaxes(a, 0, 1)  # (k, j, i) -> (j, k, i)
    scale = np.abs(uxz).max()
    assert scale > 1e-5  # the body force has moved the fluid
    for a, b in ((T(uxy), uxz), (T(uyy), uzz), (T(uzy), uyz)):
        err = np.abs(a - b).max() / scale
        print(f"[gpu] walls y vs z {storage}: max rel diff {err:.2e}", flush=True)
        assert err < tol
    np.testing.assert_allclose(T(ry), rz, rtol=tol, atol=0)


@pytest.mark.parametrize("storage,rounding", [("f16", 1), ("bf16", 0), ("f16", 0)])
def test_layouts_agree_for_16_bit_storage(gpu, storage, rounding):
    """The 16-bit forms give the same bits in either row layout (the
    stochastic-rounding stream is keyed by the cell's reference index)."""
    case = Case("ldc", 70, 19, 13, svisc=0.02, u0=0.1)
    banner(f"engine {storage} rounding={rounding} padded vs compact rows")
    out = []
    for layout in (1, 2):
        with Engine(case, storage, rounding=rounding, layout=layout, seed=3) as eng:
            eng.step(60)
            out.append(eng.populations())
            if layout == 2:
                assert eng.info()["nx"] == case.l + 2
    assert np.array_equal(out[0], out[1])
    assert np.any(out[0] != 0)


@pytest.mark.parametrize("storage,rounding,mixed", [("f32", 0, 0), ("f64", 0, 0), ("f16", 0, 0), ("f16", 1, 0),
                                                    ("bf16", 0, 0), ("f16", 0, 1), ("f32", 0, 2)])
@pytest.mark.parametrize("name", ["inflow_long", "inflow_les"])
def test_inflow_layouts_agree(gpu, name, storage, rounding, mixed):
    """Inflow boundary: padded rows (the outflow kernel before every collide)
    and compact rows (both x faces written by the collide's halo lanes in
    the forms that hold all 19 new populations, by the outflow kernel for
    the streaming FP64 collide) give the same bits, stepped in one call or
    in several (each call starts with the outflow kernel)."""
    case = CASES[name]
    banner(f"engine inflow layouts {name} {storage} rounding={rounding} mixed={mixed}")
    out = []
    for layout, chunks in ((1, [40]), (2, [40]), (2, [1, 7, 32])):
        with Engine(case, storage, rounding=rounding, layout=layout, seed=5, mixed=mixed) as eng:
            for c in chunks:
                eng.step(c)
            out.append(eng.populations())
    assert np.isfinite(out[0]).all() and np.any(out[0] != 0)
    for o in out[1:]:
        assert np.array_equal(out[0], o)


def test_engine_rejects_bad_config(gpu):
    from bgk3d_gpu_amd import runtime as rt
    with pytest.raises(rt.BGKError):
        Engine(Case("ldc", 8, 8, 8), "f32", obstacle=1)  # obstacles need inflow/periodic
    with pytest.raises(rt.BGKError):  # walls at y: the force must lie along them
        Engine(Case("channel", 8, 8, 8, wall_axis=2, force_axis=2), "f32")
    with pytest.raises(rt.BGKError):
        Engine(Case("channel", 8, 8, 8, wall_axis=1), "f32")
    with pytest.raises(rt.BGKError):
        Engine(Case("inflow", 40, 24, 24, obstacle=True), "f32", layout=3)


def solid_sphere_mask(l, m, n, centre, radius):
    """The handle API's solid sphere (obstacle 2): cells with d^2 <= R^2,
    as a (1:l,1:m,1:n) mask and its bounding box for the oracle's bcond_obs."""
    ic, jc, kc = centre
    k, j, i = np.meshgrid(np.arange(1, n + 1), np.arange(1, m + 1), np.arange(1, l + 1), indexing="ij")
    obs = (((ic - i) ** 2 + (jc - j) ** 2 + (kc - k) ** 2) <= radius * radius).astype(np.int32)
    kk, jj, ii = np.nonzero(obs)
    box = (int(ii.min()) + 1, int(ii.max()) + 1, int(jj.min()) + 1, int(jj.max()) + 1, int(kk.min()) + 1,
           int(kk.max()) + 1)
    return obs, box


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_engine_solid_sphere_matches_oracle(gpu, storage):
    """Solid sphere (the configured sphere-flow geometry; the reference builds
    only a hollow shell, SRC/build.obs.F90:52-83): the oracle's bcond_obs run
    on the same solid mask gives bit-identical fluid cells, and the drag and
    lift agree within the storage-kind accumulation of the oracle."""
    case = Case("inflow", 40, 24, 24, svisc=0.02, u0=0.1, obstacle=True)
    centre, radius = (2 * case.l // 5, case.m // 2, case.n // 2), 5
    steps = 60
    banner(f"engine solid sphere {storage} r={radius}")
    with Engine(case, storage, obstacle=2, centre=centre, radius=radius) as eng:
        eng.step(steps)
        G = eng.populations()
        dl = eng.draglift()
    prec = PREC[storage]
    obs, box = solid_sphere_mask(case.l, case.m, case.n, centre, radius)
    A = O.new_lattice(case.l, case.m, case.n, prec)
    B = O.new_lattice(case.l, case.m, case.n, prec)
    O.init(A, case.l, case.m, case.n, prec, mode=0, u00=case.u00)
    O.run("inflow", A, B, case.l, case.m, case.n, steps, prec=prec, svisc=case.svisc, u0=case.u0, obs=obs, box=box)
    fluid = obs == 0
    oi = A[:, 1:-1, 1:-1, 1:-1].astype(np.float64)
    assert np.array_equal(G[:, fluid], oi[:, fluid]), f"max |diff| {np.abs(G[:, fluid] - oi[:, fluid]).max()}"
    ref = O.draglift(A, obs, case.l, case.m, case.n, radius, prec=prec)
    assert abs(ref[0]) > 0
    tol = 1e-4 if storage == "f32" else 1e-10  # oracle sums in the storage kind, engine in FP64
    np.testing.assert_allclose(dl, ref, rtol=tol, atol=tol * np.abs(ref).max())


def test_engine_refuses_sphere_at_periodic_seam(gpu):
    from bgk3d_gpu_amd import runtime as rt
    case = Case("periodic", 40, 24, 24, svisc=0.01, u0=0.05, init=1, obstacle=True)
    with pytest.raises(rt.BGKError):
        Engine(case, "f32", centre=(16, 4, 12), radius=3)   # y box 4 +- 7 leaves 1..24
    with pytest.raises(rt.BGKError):
        Engine(case, "f32", centre=(37, 12, 12), radius=3)  # x is periodic here too
    with Engine(case, "f32", centre=(16, 12, 12), radius=3) as eng:  # the default placement is accepted
        eng.step(1)
