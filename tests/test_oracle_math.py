"""Pins of the oracle that do not rely on the reference: closed forms and
invariants chosen so that a dropped term, a wrong sign, a wrong index or a
transposed operand fails at least one of them.
"""
import numpy as np
import pytest

import oracle_support as O

CX, CY, CZ = np.array(O.CX), np.array(O.CY), np.array(O.CZ)
# weights: 1/3 rest (direction 19), 1/18 axis (5, 6, 8, 14, 15, 17), 1/36 diagonal
W = np.full(19, 1 / 36)
W[[4, 5, 7, 13, 14, 16]] = 1 / 18
W[18] = 1 / 3


def feq_shifted(rho, u):
    """Textbook D3Q19 equilibrium minus the rest weights, float64."""
    cu = CX * u[0] + CY * u[1] + CZ * u[2]
    usq = float(np.dot(u, u))
    return W * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * usq) - W


def test_lattice_tables():
    # weights sum to one, opposite pairs, isotropy (sum w c_a c_b = delta/3)
    assert abs(W.sum() - 1) < 1e-15
    opp = [11, 12, 9, 10, 13, 14, 15, 16, 17, 2, 3, 0, 1, 4, 5, 6, 7, 8, 18]
    for d in range(19):
        assert (CX[opp[d]], CY[opp[d]], CZ[opp[d]]) == (-CX[d], -CY[d], -CZ[d])
    C = np.stack([CX, CY, CZ])
    np.testing.assert_allclose((W * C) @ C.T, np.eye(3) / 3, atol=1e-15)


def uniform_state(l, m, n, rho, u, prec="dp"):
    A = O.new_lattice(l, m, n, prec)
    f = feq_shifted(rho, np.asarray(u, float))
    for d in range(19):
        A[d] = f[d]
    return A


def test_equilibrium_is_fixed_point():
    l = m = n = 6
    rho, u = 1.02, np.array([0.03, -0.02, 0.01])
    A = uniform_state(l, m, n, rho, u)
    B = O.new_lattice(l, m, n, "dp")
    O.bcond("periodic", A, l, m, n, "dp")
    O.col_mc(A, B, l, m, n, 1.3, "dp")
    O.swap(A, B)
    f = feq_shifted(rho, u)
    for d in range(19):
        np.testing.assert_allclose(A[d, 1:-1, 1:-1, 1:-1], f[d], rtol=0, atol=2e-16)


def test_omega_one_gives_closed_form_equilibrium():
    # with omega = 1 the post-collision state is e_i of the pulled moments
    rng = np.random.default_rng(1)
    l = m = n = 5
    A = O.new_lattice(l, m, n, "dp")
    A[...] = rng.standard_normal(A.shape) * 0.01
    O.bcond("periodic", A, l, m, n, "dp")
    pulled = np.stack([A[d, 1 - CZ[d]:n + 1 - CZ[d], 1 - CY[d]:m + 1 - CY[d], 1 - CX[d]:l + 1 - CX[d]]
                       for d in range(19)])
    B = O.new_lattice(l, m, n, "dp")
    O.col_mc(A, B, l, m, n, 1.0, "dp")
    O.swap(A, B)
    rho = pulled.sum(axis=0) + 1
    u = np.stack([np.tensordot(c, pulled, axes=(0, 0)) for c in (CX, CY, CZ)]) / rho
    for idx in [(0, 0, 0), (2, 3, 1), (4, 4, 4)]:
        f = feq_shifted(rho[idx], u[(slice(None),) + idx])
        np.testing.assert_allclose(A[(slice(None),) + tuple(i + 1 for i in idx)], f, atol=1e-15)


@pytest.mark.parametrize("prec", ["dp", "sp"])
def test_collide_conserves_mass_and_momentum(prec):
    rng = np.random.default_rng(2)
    l, m, n = 7, 6, 5
    A = O.new_lattice(l, m, n, prec)
    A[...] = rng.standard_normal(A.shape) * 0.01
    O.bcond("periodic", A, l, m, n, prec)
    I0 = A[:, 1:-1, 1:-1, 1:-1].astype(np.float64)
    mass0 = I0.sum()
    mom0 = [np.tensordot(c, I0, axes=(0, 0)).sum() for c in (CX, CY, CZ)]
    B = O.new_lattice(l, m, n, prec)
    O.col_mc(A, B, l, m, n, 1.7, prec)
    O.swap(A, B)
    I1 = A[:, 1:-1, 1:-1, 1:-1].astype(np.float64)
    tol = 1e-13 if prec == "dp" else 2e-5
    assert abs(I1.sum() - mass0) < tol
    for c, p0 in zip((CX, CY, CZ), mom0):
        assert abs(np.tensordot(c, I1, axes=(0, 0)).sum() - p0) < tol


def test_pure_streaming_is_a_permutation():
    # omega = 0: L periodic steps on an L^3 box return the state bitwise,
    # and one step moves every population by c_d
    rng = np.random.default_rng(3)
    L = 6
    A = O.new_lattice(L, L, L, "dp")
    A[...] = rng.standard_normal(A.shape)
    start = A[:, 1:-1, 1:-1, 1:-1].copy()
    B = O.new_lattice(L, L, L, "dp")
    O.bcond("periodic", A, L, L, L, "dp")
    O.col_mc(A, B, L, L, L, 0.0, "dp")
    O.swap(A, B)
    for d in range(19):
        np.testing.assert_array_equal(A[d, 1:-1, 1:-1, 1:-1], np.roll(start[d], (CZ[d], CY[d], CX[d]), (0, 1, 2)))
    for _ in range(L - 1):
        O.bcond("periodic", A, L, L, L, "dp")
        O.col_mc(A, B, L, L, L, 0.0, "dp")
        O.swap(A, B)
    np.testing.assert_array_equal(A[:, 1:-1, 1:-1, 1:-1], start)


def test_body_force_adds_ten_fgrad_rho_momentum():
    rng = np.random.default_rng(4)
    l, m, n = 5, 4, 4
    A = O.new_lattice(l, m, n, "dp")
    A[...] = rng.standard_normal(A.shape) * 0.01
    O.bcond("periodic", A, l, m, n, "dp")
    pulled = np.stack([A[d, 1 - CZ[d]:n + 1 - CZ[d], 1 - CY[d]:m + 1 - CY[d], 1 - CX[d]:l + 1 - CX[d]]
                       for d in range(19)])
    rho = pulled.sum(axis=0) + 1
    B0, B1 = O.new_lattice(l, m, n, "dp"), O.new_lattice(l, m, n, "dp")
    A1 = A.copy()
    O.col_mc(A, B0, l, m, n, 1.1, "dp")
    O.col_mc(A1, B1, l, m, n, 1.1, "dp", force_mode=1, fgrad=1e-3)
    dmx = sum(CX[d] * (B1[d] - B0[d]) for d in range(18))[1:-1, 1:-1, 1:-1]
    dmass = sum(B1[d] - B0[d] for d in range(18))[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(dmx, 10 * 1e-3 * rho, rtol=1e-12)
    np.testing.assert_allclose(dmass, 0, atol=1e-15)


def test_driven_cavity_mirror_symmetry():
    # the lid moves along x, so the flow is symmetric under y -> m+1-y:
    # u_x, u_z even, u_y odd
    l, m, n = 10, 9, 8
    A, B = O.new_lattice(l, m, n, "dp"), O.new_lattice(l, m, n, "dp")
    O.init(A, l, m, n, "dp")
    O.run("ldc", A, B, l, m, n, 60, prec="dp", svisc=0.05, u0=0.1)
    I = A[:, 1:-1, 1:-1, 1:-1]
    rho = I.sum(axis=0) + 1
    u = [np.tensordot(c, I, axes=(0, 0)) / rho for c in (CX, CY, CZ)]
    np.testing.assert_allclose(u[0], u[0][:, ::-1, :], atol=1e-15)
    np.testing.assert_allclose(u[2], u[2][:, ::-1, :], atol=1e-15)
    np.testing.assert_allclose(u[1], -u[1][:, ::-1, :], atol=1e-15)
    assert np.abs(u[0]).max() > 1e-3
    # total mass is conserved by bounce-back walls and the lid
    assert abs((rho - 1).sum()) < 1e-12


def test_taylor_green_energy_decay():
    # 3-D Taylor-Green kinetic energy decays as exp(-6 nu k0^2 t) in the
    # linear regime; the lattice error must fall at second order in 1/L
    errs = []
    for L, steps in ((24, 200), (32, 356)):  # same physical time t nu k0^2
        nu, u0 = 0.02, 0.001
        A, B = O.new_lattice(L, L, L, "dp"), O.new_lattice(L, L, L, "dp")
        O.init(A, L, L, L, "dp", mode=1, u00=u0)
        e0 = O.diagno(A, L, L, L, "dp")[4]
        O.run("periodic", A, B, L, L, L, steps, prec="dp", svisc=nu, u0=u0, nthreads=8)
        e1 = O.diagno(A, L, L, L, "dp")[4]
        k0 = 2 * np.pi / L
        errs.append(e1 / e0 / np.exp(-6 * nu * k0 ** 2 * steps) - 1)
    assert abs(errs[0]) < 0.1 and abs(errs[1]) < 0.06
    ratio = errs[1] / errs[0]
    assert 0.45 < ratio < 0.70, errs  # (24/32)^2 = 0.5625


def test_poiseuille_profile():
    # channel with half-way bounce-back walls at z = 1/2 and n + 1/2 and the
    # body force fgrad*rho on each of the 10 populations with c_x = +-1
    # (momentum g = 10 fgrad per step).  The steady profile is a parabola of
    # curvature -g/nu; the stored (post-collision) momentum carries +g/2 over
    # the mid-step velocity, plus the small BGK wall slip.
    l, m, n, nu, u0 = 4, 4, 12, 1.0 / 6.0, 0.005
    A, B = O.new_lattice(l, m, n, "dp"), O.new_lattice(l, m, n, "dp")
    O.init(A, l, m, n, "dp")
    O.run("channel", A, B, l, m, n, 6000, prec="dp", svisc=nu, u0=u0)
    _, fgrad = O.hencol(nu, u0, m, "dp")
    g = 10 * fgrad
    I = A[:, 1:-1, 1:-1, 1:-1]
    rho = I.sum(axis=0) + 1
    ux = (np.tensordot(CX, I, axes=(0, 0)) / rho).mean(axis=(1, 2))
    np.testing.assert_allclose(ux[2:] - 2 * ux[1:-1] + ux[:-2], -g / nu, rtol=1e-6)
    np.testing.assert_allclose(ux, ux[::-1], rtol=1e-12)
    z = np.arange(1, n + 1)
    offset = ux - g / (2 * nu) * (z - 0.5) * (n + 0.5 - z)
    np.testing.assert_allclose(offset, offset.mean(), rtol=1e-6)
    assert abs(offset.mean() - g / 2) < 0.15 * g


def test_reductions_on_uniform_state():
    l, m, n = 6, 5, 4
    rho, u = 1.01, np.array([0.02, -0.01, 0.005])
    A = uniform_state(l, m, n, rho, u)
    d = O.diagno(A, l, m, n, "dp")
    np.testing.assert_allclose(d[:3], u / 1.0, rtol=1e-12)
    np.testing.assert_allclose(d[3], rho, rtol=1e-14)
    np.testing.assert_allclose(d[4], 0.5 * np.dot(u, u), rtol=1e-12)
    uk, wk, vk, den = O.varm(A, l, m, n, "dp")
    # per-plane sums of l*m (resp. l*n) equal cells, times 1/l in default real
    np.testing.assert_allclose(uk, m * u[0], rtol=1e-6)
    np.testing.assert_allclose(wk, m * u[1], rtol=1e-6)
    np.testing.assert_allclose(vk, m * u[2], rtol=1e-6)
    np.testing.assert_allclose(den, n * rho, rtol=1e-6)
    pk = O.prof_k(A, l, m, n, 2, 2, "dp")
    np.testing.assert_allclose(pk[:, 0], u[0], rtol=1e-12)
    np.testing.assert_allclose(pk[:, 1], u[1], rtol=1e-12)
    np.testing.assert_allclose(pk[:, 2], u[2], rtol=1e-12)
    np.testing.assert_allclose(pk[:, 3], rho, rtol=1e-14)


def test_draglift_brute_force():
    # momentum exchange sum_d 2 c_d f_d(x) obs(x + c_d) over fluid cells of
    # the box, by direct numpy enumeration
    rng = np.random.default_rng(5)
    l, m, n = 40, 24, 24
    obs, box, radius = O.build_obs(l, m, n)
    A = O.new_lattice(l, m, n, "dp")
    A[...] = rng.standard_normal(A.shape) * 0.01
    got = O.draglift(A, obs, l, m, n, radius, u_inflow=0.1, prec="dp")
    ic, jc, kc = 2 * l // 5, m // 2, n // 2
    F = np.zeros(3)
    for k in range(kc - radius - 3, kc + radius + 4):
        for j in range(jc - radius - 3, jc + radius + 4):
            for i in range(ic - radius - 3, ic + radius + 4):
                if obs[k - 1, j - 1, i - 1]:
                    continue
                for d in range(18):
                    if obs[k + CZ[d] - 1, j + CY[d] - 1, i + CX[d] - 1]:
                        F += 2 * np.array([CX[d], CY[d], CZ[d]]) * A[d, k, j, i]
    np.testing.assert_allclose(got, F / (0.1 * 0.1 * radius), rtol=1e-12, atol=1e-14)


def test_build_obs_shell():
    obs, box, radius = O.build_obs(100, 48, 48)
    assert radius == 6
    k, j, i = np.nonzero(obs)
    d2 = (i + 1 - 40) ** 2 + (j + 1 - 24) ** 2 + (k + 1 - 24) ** 2
    assert d2.min() > 16 and d2.max() < 64
    assert box == (int(i.min()) + 1, int(i.max()) + 1, int(j.min()) + 1, int(j.max()) + 1, int(k.min()) + 1,
                   int(k.max()) + 1)


def test_taylor_green_init_with_density_field():
    """Init mode 2 keeps the density perturbation: the zeroth moment of the
    stored state is 1 + u0^2/48 (cos 4 pi x + cos 4 pi y)(cos 4 pi z + 2) and
    the velocity the Taylor-Green field; mode 1 (the reference) has density
    1 everywhere and the same momentum times the perturbation."""
    l = m = n = 16
    u0 = 0.05
    A1, A2 = O.new_lattice(l, m, n, "dp"), O.new_lattice(l, m, n, "dp")
    O.init(A1, l, m, n, "dp", mode=1, u00=u0)
    O.init(A2, l, m, n, "dp", mode=2, u00=u0)
    pi = float(np.float32(np.pi))       # default-real pi (SRC/init.F90:63)
    x = (np.arange(0, l + 2) - 0.5) / l
    X = x[None, None, :]
    Y = x[None, :, None]
    Z = x[:, None, None]
    rho_exact = 1 + u0 ** 2 / 48 * (np.cos(4 * pi * X) + np.cos(4 * pi * Y)) * (np.cos(4 * pi * Z) + 2)
    ux = u0 * np.sin(2 * pi * X) * np.cos(2 * pi * Y) * np.cos(2 * pi * Z)
    uy = -u0 * np.cos(2 * pi * X) * np.sin(2 * pi * Y) * np.cos(2 * pi * Z)
    cx, cy, cz = (np.array(c, float)[:, None, None, None] for c in (O.CX, O.CY, O.CZ))
    for A, rho_expect in ((A2, rho_exact), (A1, np.ones_like(rho_exact))):
        rho = A.sum(axis=0) + 1.0
        np.testing.assert_allclose(rho, np.broadcast_to(rho_expect, rho.shape), rtol=0, atol=1e-15)
        jx, jy, jz = (cx * A).sum(axis=0), (cy * A).sum(axis=0), (cz * A).sum(axis=0)
        # momentum = crho * u in both forms
        np.testing.assert_allclose(jx, rho_exact * ux, rtol=0, atol=1e-15)
        np.testing.assert_allclose(jy, rho_exact * uy, rtol=0, atol=1e-15)
        np.testing.assert_allclose(jz, 0.0, atol=1e-15)
