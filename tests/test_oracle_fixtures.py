"""The C oracle against golden fixtures produced by the reference's own
Fortran routines (tests/golden/make_fixtures.py).

Bit-exact: the sha256 of the oracle's interior populations equals the
reference's for every run (FP32 and FP64), and every single-routine pin
(one `boundaries` or one `col_MC` call on a seeded random state) matches.
The reference's text diagnostics (diagno.dat, probe.dat, prof_i.dat,
prof_j.dat, prof_k.dat, u_med.dat) are reproduced character for character
from the oracle's values, and the files its writers produce (restart,
velocity, VTK planes) byte for byte by the numpy restatement of the formats
applied to the oracle's state (restart: the interior of every record).
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import format_support as F
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
    texts = {"diagno": [], "prof_i": {}, "prof_j": {}, "prof_k": {}, "probe": [], "drag": {}}
    step_to = 0
    # SRC/diagnostic.F90:98-120
    pt = (l // 3, m // 5, n // 3) if spec["case"] == "periodic" else (l // 2, m // 2, n // 2)

    def sample(t):
        texts["diagno"].append(O.diagno_line(t, O.diagno(A, l, m, n, prec)))
        texts["probe"].append(f"{t:8d}" + "".join(O.fortran_e(v) + " " for v in O.probe(A, *pt, prec)))
        texts["prof_i"][t] = O.prof_line(A, l, m, n, 0, m // 2, n // 2, prec)
        texts["prof_j"][t] = O.prof_line(A, l, m, n, 1, l // 2, n // 2, prec)
        texts["prof_k"][t] = O.prof_k(A, l, m, n, l // 2, m // 2, prec)
        if obs is not None and t > 0:
            texts["drag"][t] = O.draglift(A, obs, l, m, n, radius, prec=prec)

    sample(0)
    u_med0 = O.varm(A, l, m, n, prec)
    while step_to < spec["steps"]:
        chunk = min(spec["icheck"], spec["steps"] - step_to)
        O.run(spec["case"], A, B, l, m, n, chunk, prec=prec, svisc=spec["svisc"], u0=spec["u0"], obs=obs, box=box,
              nthreads=threads, les=spec.get("les", False))
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
    if "sha256_interior_all" in spec:
        # solid cells too: the oracle's bcond_obs runs the reference's serial loop order
        assert digest(A[:, 1:-1, 1:-1, 1:-1]) == spec["sha256_interior_all"]
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
    # probe.dat (i8, 4e14.6), line for line
    ref_lines = [x for x in spec["texts"]["probe.dat"].splitlines() if x.strip()]
    assert len(ref_lines) == len(texts["probe"])
    for a, b in zip(ref_lines, texts["probe"]):
        assert a.rstrip() == b.rstrip(), (a, b)
    # prof_{i,j,k}.dat blocks: (q-0.5, u, w, v, den) in e14.6
    for name in ("prof_i", "prof_j", "prof_k"):
        blocks = O.parse_blocks(spec["texts"][f"{name}.dat"])
        assert set(blocks) == set(texts[name])
        for t, rows in blocks.items():
            pk = texts[name][t]
            assert len(rows) == len(pk)
            for q, row in enumerate(rows):
                expect = "".join(O.fortran_e(v) + " " for v in (q + 0.5, *pk[q]))
                assert row.rstrip() == expect.rstrip(), (name, t, q, row, expect)
    # files the reference's writers produced at the last step
    if "files" not in spec:
        return
    files = spec["files"]
    steps = spec["steps"]
    l, m, n, prec = spec["l"], spec["m"], spec["n"], spec["prec"]
    ours = {
        f"save.{steps:08d}.bin": F.restart_bytes(steps, A, l, m, n),
        f"save_vel.{steps:08d}.bin": F.save_vel_bytes(steps, A, l, m, n, prec),
        f"tec_xy.{steps:08d}.vtk": F.vtk_bytes(0, l, m, n, n // 2, F.vtk_plane_values(A, l, m, n, 0, n // 2, prec)),
        f"tec_xz.{steps:08d}.vtk": F.vtk_bytes(1, l, m, n, m // 2, F.vtk_plane_values(A, l, m, n, 1, m // 2, prec)),
        f"tec_yz.{steps:08d}.vtk": F.vtk_bytes(2, l, m, n, l // 2, F.vtk_plane_values(A, l, m, n, 2, l // 2, prec)),
        f"tec_om.xy.{steps:08d}.vtk": F.vtk_bytes(3, l, m, n, n // 2, F.vorticity_xy_values(A, l, m, n, n // 2, prec)),
    }
    got = F.file_digests(ours, l, m, n, prec)
    assert set(got) == set(files)
    for key in files:
        if key == "save_raw" and obs is not None:
            # solid cells are not part of the parity claim (obstacle deviation)
            assert got[key]["itime"] == files[key]["itime"]
            assert got[key]["record_bytes"] == files[key]["record_bytes"]
            continue
        if obs is not None and key != "save_raw":
            continue
        assert got[key] == files[key], key


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


NOSHIFT_KEYS = sorted(FIX["noshift"])


@pytest.mark.parametrize("key", NOSHIFT_KEYS)
def test_oracle_reproduces_reference_noshift_run(key):
    """The reference's -DNOSHIFT builds (populations stored unshifted,
    SRC/col_MC.F90:139-144, SRC/init.F90:71-75, SRC/bcond.inflow.F90:68-72,
    SRC/diagno.F90:55-59): every interior cell and the text diagnostics."""
    spec = FIX["noshift"][key]
    with O.noshift():
        A, obs, texts, _, _ = run_oracle(spec)
    assert digest(A[:, 1:-1, 1:-1, 1:-1]) == spec["sha256_interior"]
    for name, ours in (("diagno.dat", texts["diagno"]), ("probe.dat", texts["probe"])):
        ref_lines = [x.rstrip() for x in spec["texts"][name].splitlines() if x.strip()]
        assert ref_lines == [x.rstrip() for x in ours], name
    blocks = O.parse_blocks(spec["texts"]["prof_k.dat"])
    for t, rows in blocks.items():
        pk = texts["prof_k"][t]
        for q, row in enumerate(rows):
            expect = "".join(O.fortran_e(v) + " " for v in (q + 0.5, *pk[q]))
            assert row.rstrip() == expect.rstrip(), (t, q)


def test_noshift_state_is_shifted_state_plus_weights():
    """Unshifted and shifted runs of the same cavity hold the same physics:
    after 50 steps the unshifted populations minus the weights equal the
    shifted ones to rounding (the two differ only in where w_i is kept)."""
    l = m = n = 12
    res = {}
    for flag in (False, True):
        A = O.new_lattice(l, m, n, "dp")
        B = O.new_lattice(l, m, n, "dp")
        ctx = O.noshift() if flag else None
        if ctx:
            ctx.__enter__()
        try:
            O.init(A, l, m, n, "dp", mode=0, u00=0.1)
            O.run("ldc", A, B, l, m, n, 50, prec="dp", svisc=0.05, u0=0.1)
        finally:
            if ctx:
                ctx.__exit__(None, None, None)
        res[flag] = A[:, 1:-1, 1:-1, 1:-1].copy()
    w = np.array([1 / 36] * 4 + [1 / 18] * 2 + [1 / 36] + [1 / 18] + [1 / 36] * 5 + [1 / 18] * 2 + [1 / 36]
                 + [1 / 18] + [1 / 36] + [1 / 3])
    np.testing.assert_allclose(res[True] - w[:, None, None, None], res[False], atol=1e-14, rtol=0)
    assert np.abs(res[False]).max() > 1e-4


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


RESTART_KEYS = sorted(FIX["restart"])


@pytest.mark.parametrize("key", RESTART_KEYS)
def test_restart_file_of_oracle_state(key):
    """The oracle's state at the save step, written in the restart format,
    has the interior records of the reference's own save_raw file."""
    r = FIX["restart"][key]
    spec = dict(FIX["runs"][r["run"]])
    spec["steps"] = r["save_step"]
    spec["icheck"] = r["save_step"]
    A, *_ = run_oracle(spec)
    l, m, n, prec = spec["l"], spec["m"], spec["n"], spec["prec"]
    got = F.restart_interior_digest(F.restart_bytes(r["save_step"], A, l, m, n), l, m, n, prec)
    assert got == r["save_raw"]


@pytest.mark.parametrize("key", RESTART_KEYS)
def test_reference_restores_our_restart_file(key):
    """The reference's restore_raw reads a file written from the oracle's
    state and its run from there ends on the full run's state."""
    r = FIX["restart"][key]
    spec = FIX["runs"][r["run"]]
    if not O.ref_available(spec["case"], spec["prec"]):
        pytest.skip("reference build not available")
    half = dict(spec)
    half["steps"] = half["icheck"] = r["save_step"]
    A, *_ = run_oracle(half)
    l, m, n, prec = spec["l"], spec["m"], spec["n"], spec["prec"]
    B, _, _ = O.run_ref(spec["case"], l, m, n, spec["steps"], prec=prec, svisc=spec["svisc"], u0=spec["u0"],
                        restore=F.restart_bytes(r["save_step"], A, l, m, n))
    assert digest(np.ascontiguousarray(B[:, 1:-1, 1:-1, 1:-1])) == spec["sha256_interior"]


ORIG_KEYS = sorted(FIX["original"])


@pytest.mark.parametrize("key", ORIG_KEYS)
def test_original_two_pass_build_equals_fused(key):
    """The reference's ORIGINAL build (boundaries, movef, col: SRC/bgk3d.F90:177-179)
    ends on exactly the fused build's state, so the fused kernel stands in for
    both; recorded by the fixture generator and re-run here when the
    reference build is present."""
    rec = FIX["original"][key]
    spec = FIX["runs"][key]
    assert rec["equals_fused"] and rec["sha256_interior"] == spec["sha256_interior"]
    if not O.ref_available(spec["case"], spec["prec"], variant="_orig"):
        pytest.skip("reference ORIGINAL build not available")
    A, _, _ = O.run_ref(spec["case"], spec["l"], spec["m"], spec["n"], spec["steps"], prec=spec["prec"],
                        svisc=spec["svisc"], u0=spec["u0"], icheck=spec["icheck"], init_v=spec["init"],
                        variant="_orig")
    assert digest(np.ascontiguousarray(A[:, 1:-1, 1:-1, 1:-1])) == spec["sha256_interior"]


@pytest.mark.parametrize("case,key", [("periodic", "tgv_32_sp"), ("ldc", "ldc_24_sp")])
def test_threaded_reference_build_reproduces_serial_run(case, key):
    """The multicore CPU baseline: the reference's OpenMP loops (-DOFFLOAD
    branches, with their private lists) built without an offload target and
    run on 4 host threads end on the serial build's fixture state bit for
    bit.  (Its `do concurrent` build mapped to OpenMP, -fdo-concurrent-to-
    openmp=host, shares the loop temporaries of SRC/col_MC.F90:272-519
    between threads and does not: that is why bench.py times this one.)"""
    spec = FIX["runs"][key]
    if not O.ref_available(case, "sp", variant="_omphost"):
        pytest.skip("threaded reference build not available")
    A, texts, _ = O.run_ref(case, spec["l"], spec["m"], spec["n"], spec["steps"], prec="sp", svisc=spec["svisc"],
                            u0=spec["u0"], icheck=spec["icheck"], init_v=spec["init"], threads=4,
                            variant="_omphost")
    assert digest(A[:, 1:-1, 1:-1, 1:-1]) == spec["sha256_interior"]
    ref_lines = [x.rstrip() for x in spec["texts"]["diagno.dat"].splitlines() if x.strip()]
    assert [x.rstrip() for x in texts["diagno.dat"].splitlines() if x.strip()] == ref_lines
