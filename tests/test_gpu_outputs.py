"""Diagnostics and output files on the GPU against the reference's own.

* v1 ABI: prof_i/prof_j/prof_k/probe bit-identical to the oracle on seeded
  random states (FP32, FP64), padded layout.
* Handle API, on the golden runs: probe.dat and prof_{i,j,k}.dat values at
  the last step printed as the reference prints them equal its lines; the
  restart file (interior records), the velocity file and the four VTK files
  equal the digests of the files the reference's writers produced.
* Restart: a handle restored from a file written from the oracle's state at
  half the steps (the reference's format) and run on ends on the reference's
  full-run state; so does one restored from its own file.
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np
import pytest

import format_support as F
import oracle_support as O
from bgk3d_gpu_amd.cases import Case
from bgk3d_gpu_amd.engine import Engine
from bgk3d_gpu_amd.simulation import Simulation

pytestmark = pytest.mark.gpu

FIX = json.loads((Path(__file__).parent / "golden" / "fixtures.json").read_text())
PLAIN_RUNS = ["ldc_24_sp", "ldc_24_dp", "tgv_32_sp", "tgv_32_dp", "channel_64x32x32_sp", "channel_64x32x32_dp"]


def banner(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def case_of(spec):
    return Case(spec["case"], spec["l"], spec["m"], spec["n"], svisc=spec["svisc"], u0=spec["u0"], init=spec["init"])


@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_v1_line_profiles_and_probe_match_oracle(gpu, prec):
    l, m, n = 19, 13, 11
    banner(f"v1 profiles {prec} {l}x{m}x{n}")
    case = Case("periodic", l, m, n, ipad=5, jpad=2)
    sim = Simulation(case, prec)
    rng = np.random.default_rng(5)
    A = O.new_lattice(l, m, n, prec, ipad=5, jpad=2)
    A[...] = (rng.standard_normal(A.shape) * 0.01).astype(A.dtype)
    sim.set_state(A)
    for axis, (c1, c2) in ((0, (m // 2, n // 2)), (1, (l // 2, n // 2)), (2, (l // 2, m // 2)), (0, (0, n + 1))):
        ref = O.prof_line(A, l, m, n, axis, c1, c2, prec)
        got = {0: sim.prof_i, 1: sim.prof_j, 2: sim.prof_k}[axis](c1, c2)
        assert np.array_equal(got, ref), (axis, c1, c2)
    for pt in ((1, 1, 1), (l, m, n), (l // 3, m // 5, n // 3)):
        assert np.array_equal(sim.probe(*pt), O.probe(A, *pt, prec)), pt


@pytest.mark.parametrize("key", PLAIN_RUNS)
def test_engine_outputs_match_reference_files(gpu, key, tmp_path):
    spec = FIX["runs"][key]
    l, m, n, prec, steps = spec["l"], spec["m"], spec["n"], spec["prec"], spec["steps"]
    banner(f"engine outputs {key}")
    with Engine(case_of(spec), "f32" if prec == "sp" else "f64") as eng:
        eng.step(steps)
        # last probe.dat line and the last profile blocks
        pt = (l // 3, m // 5, n // 3) if spec["case"] == "periodic" else (l // 2, m // 2, n // 2)
        line = f"{steps:8d}" + "".join(O.fortran_e(v) + " " for v in eng.probe(*pt))
        assert spec["texts"]["probe.dat"].splitlines()[-1].rstrip() == line.rstrip()
        for name, axis, c in (("prof_i", 0, (m // 2, n // 2)), ("prof_j", 1, (l // 2, n // 2)),
                              ("prof_k", 2, (l // 2, m // 2))):
            rows = O.parse_blocks(spec["texts"][f"{name}.dat"])[steps]
            vals = eng.profile(axis, *c)
            for q, row in enumerate(rows):
                expect = "".join(O.fortran_e(v) + " " for v in (q + 0.5, *vals[q]))
                assert row.rstrip() == expect.rstrip(), (name, q)
        # files
        eng.save_raw(tmp_path / "save.bin", steps)
        eng.save_vel(tmp_path / "save_vel.bin", steps)
        eng.write_vtk(0, n // 2, tmp_path / "xy.vtk")
        eng.write_vtk(1, m // 2, tmp_path / "xz.vtk")
        eng.write_vtk(2, l // 2, tmp_path / "yz.vtk")
        eng.write_vtk(3, n // 2, tmp_path / "om.vtk")
    files = spec["files"]
    assert F.restart_interior_digest((tmp_path / "save.bin").read_bytes(), l, m, n, prec) == files["save_raw"]
    assert F.sha((tmp_path / "save_vel.bin").read_bytes()) == files["save_vel"]["sha256"]
    for name, key2 in (("xy", "vtk_xy"), ("xz", "vtk_xz"), ("yz", "vtk_yz"), ("om", "vtk_om_xy")):
        b = (tmp_path / f"{name}.vtk").read_bytes()
        assert F.vtk_header(b) == files[key2]["header"], name
        assert F.sha(b) == files[key2]["sha256"], name


@pytest.mark.parametrize("key", sorted(FIX["restart"]))
def test_engine_restart_round_trips(gpu, key, tmp_path):
    r = FIX["restart"][key]
    spec = FIX["runs"][r["run"]]
    l, m, n, prec, half = spec["l"], spec["m"], spec["n"], spec["prec"], r["save_step"]
    storage = "f32" if prec == "sp" else "f64"
    banner(f"restart {key}")
    # a reference-format file from the oracle's state at the save step
    A = O.new_lattice(l, m, n, prec)
    B = O.new_lattice(l, m, n, prec)
    O.init(A, l, m, n, prec, mode=spec["init"], u00=spec["u0"])
    O.run(spec["case"], A, B, l, m, n, half, prec=prec, svisc=spec["svisc"], u0=spec["u0"], nthreads=8)
    ref_file = tmp_path / "restore.bin"
    ref_file.write_bytes(F.restart_bytes(half, A, l, m, n))
    with Engine(case_of(spec), storage) as eng:
        assert eng.restore_raw(ref_file) == half
        eng.step(spec["steps"] - half)
        assert digest(eng.populations().astype(F.KIND[prec])) == spec["sha256_interior"]
    # its own file
    with Engine(case_of(spec), storage) as eng:
        eng.step(half)
        eng.save_raw(tmp_path / "own.bin", half)
        assert F.restart_interior_digest((tmp_path / "own.bin").read_bytes(), l, m, n, prec) == r["save_raw"]
    with Engine(case_of(spec), storage) as eng:
        assert eng.restore_raw(tmp_path / "own.bin") == half
        eng.step(spec["steps"] - half)
        assert digest(eng.populations().astype(F.KIND[prec])) == spec["sha256_interior"]


def test_restore_rejects_a_wrong_grid(gpu, tmp_path):
    with Engine(Case("ldc", 12, 10, 8), "f32") as eng:
        eng.step(3)
        eng.save_raw(tmp_path / "s.bin", 3)
    with Engine(Case("ldc", 12, 10, 9), "f32") as eng:
        with pytest.raises(Exception):
            eng.restore_raw(tmp_path / "s.bin")


@pytest.mark.parametrize("storage,prec", [("f32", "sp"), ("f64", "dp")])
def test_engine_taylor_green_init_with_density_field(gpu, storage, prec):
    l, m, n = 24, 20, 16
    banner(f"init mode 2 {storage}")
    A = O.new_lattice(l, m, n, prec)
    O.init(A, l, m, n, prec, mode=2, u00=0.05)
    with Engine(Case("periodic", l, m, n, svisc=0.01, u0=0.05, init=2), storage) as eng:
        P = eng.populations()
    assert np.array_equal(P, A[:, 1:-1, 1:-1, 1:-1].astype(np.float64))
