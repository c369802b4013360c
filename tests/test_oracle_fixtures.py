"""The C oracle against golden fixtures produced by the reference's own
Fortran routines (tests/golden/make_fixtures.py).

Bit-exact: the sha256 of the oracle's interior populations equals the
reference's for every run (FP32 and FP64), and every single-routine pin
(one `boundaries` or one `col_MC` call on a seeded random state) matches.
The reference's text diagnostics (diagno.dat, prof_k.dat, u_med.dat) are
reproduced character for character from the oracle's values.
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import oracle_support as O

FIX = json.loads((Path(__file__).parent / "golden" / "fixtures.json").read_text())


def digest(arr):
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def run_oracle(spec, threads=8):
    l, m, n, prec = spec["l"], spec["m"], spec["n"], spec["prec"]
    A = O.new_lattice(l, m, n, prec)
    B = O.new_lattice(l, m, n, prec)
    O.init(A, l, m, n, prec, mode=spec["init"], u00=spec["u0"] if spec["case"] != "channel" else 0.0)
    obs = box = None
    radius = None
    if spec["case"] == "inflow":
        obs, box, radius = O.build_obs(l, m, n)
    texts = {"diagno": [], "prof_k": {}, "drag": {}}
    step_to = 0

    def sample(t):
        texts["diagno"].append(O.diagno_line(t, O.diagno(A, l, m, n, prec)))
        texts["prof_k"][t] = O.prof_k(A, l, m, n, l // 2, m // 2, prec)
        if obs is not None and t > 0:
            texts["drag"][t] = O.draglift(A, obs, l, m, n, radius, prec=prec)

    sample(0)
    u_med0 = O.varm(A, l, m, n, prec)
    while step_to < spec["steps"]:
        chunk = min(spec["icheck"], spec["steps"] - step_to)
        O.run(spec["case"], A, B, l, m, n, chunk, prec=prec, svisc=spec["svisc"], u0=spec["u0"], obs=obs, box=box,
              nthreads=threads)
        step_to += chunk
        if step_to % spec["icheck"] == 0:
            sample(step_to)
    return A, obs, texts, u_med0, O.varm(A, l, m, n, prec)


RUN_KEYS = sorted(FIX["runs"])


@pytest.mark.parametrize("key", RUN_KEYS)
def test_oracle_reproduces_reference_run(key):
    spec = FIX["runs"][key]
    A, obs, texts, u0, u1 = run_oracle(spec)
    I = np.ascontiguousarray(A[:, 1:-1, 1:-1, 1:-1])
    if obs is not None:
        I = np.ascontiguousarray(I[:, obs == 0])
    sums = I.reshape(19, -1).astype(np.float64).sum(axis=1)
    np.testing.assert_allclose(sums, spec["pop_sums"], rtol=1e-6, atol=1e-9)
    assert digest(I) == spec["sha256_interior"], "interior populations differ from the reference"
    # diagno.dat, line for line
    ref_lines = [x for x in spec["texts"]["diagno.dat"].splitlines() if x.strip()]
    ours = texts["diagno"]
    assert len(ref_lines) == len(ours)
    for a, b in zip(ref_lines, ours):
        assert a.rstrip() == b.rstrip(), (a, b)
    # drag.lift.dat (list-directed output of the storage kind)
    for line in spec["texts"]["drag.lift.dat"].splitlines():
        if not line.strip():
            continue
        t, *vals = line.split()
        np.testing.assert_allclose(texts["drag"][int(t)], [float(v) for v in vals], rtol=2e-7 if spec["prec"] == "sp"
                                   else 1e-14, atol=0)
    # prof_k.dat blocks: (k-0.5, u, w, v, den) in e14.6
    blocks = O.parse_blocks(spec["texts"]["prof_k.dat"])
    for t, rows in blocks.items():
        pk = texts["prof_k"][t]
        for k, row in enumerate(rows):
            expect = "".join(O.fortran_e(v) + " " for v in (k + 0.5, *pk[k]))
            assert row.rstrip() == expect.rstrip(), (t, k, row, expect)


def test_oracle_reproduces_reference_varm():
    spec = FIX["runs"]["ldc_24_sp"]
    _, _, _, u0, u1 = run_oracle(spec)
    blocks = O.parse_blocks(spec["texts"]["u_med.dat"])
    # varm prints rows j = 1..m of (u(j), w(j), den(j)) * (1/l) (SRC/varm.F90:96-100)
    for t, (u, w, v, den) in ((0, u0), (spec["steps"], u1)):
        rows = blocks[t]
        for j, row in enumerate(rows):
            expect = f"{j + 1:5d}" + "".join(O.fortran_e(x) + " " for x in (u[j], w[j], den[j]))
            assert row.rstrip() == expect.rstrip(), (t, j, row, expect)


UNIT_KEYS = sorted(FIX["unit"])


@pytest.mark.parametrize("key", UNIT_KEYS)
def test_single_routine_pins(key):
    spec = FIX["unit"][key]
    l, m, n, prec, case = spec["l"], spec["m"], spec["n"], spec["prec"], spec["case"]
    rng = np.random.default_rng(spec["seed"])
    A = O.new_lattice(l, m, n, prec)
    A[...] = (rng.standard_normal(A.shape) * 0.01).astype(A.dtype)
    if spec["mode"] == 1:
        obs = box = None
        if case == "inflow":
            obs, box, _ = O.build_obs(l, m, n)
        O.bcond(case, A, l, m, n, prec, u00=spec["u0"], obs=obs, box=box)
        # every halo value the next pull reads
        assert digest(O.pulled_set(A, l, m, n)) == spec["sha256_pulled"]
        if case != "channel":
            assert digest(A) == spec["sha256_full"]
        # (the reference's CPU channel wall loop runs over 0:l+1 x 0:m+1 and
        # writes outside the arrays, SRC/bcond.channel.F90:253-267; those
        # writes land only in cells no pull reads, so the full arrays differ
        # while the pulled set above is identical)
    else:
        B = O.new_lattice(l, m, n, prec)
        omega, fgrad = O.hencol(spec["svisc"], spec["u0"] if case == "channel" else 0.0, m, prec)
        O.col_mc(A, B, l, m, n, omega, prec, force_mode=1 if case == "channel" else 0, fgrad=fgrad)
        O.swap(A, B)
        assert digest(np.ascontiguousarray(A[:, 1:-1, 1:-1, 1:-1])) == spec["sha256_interior"]
