"""CPU checks of the C-ABI libraries and the host-side logic (no GPU calls).

* every function declared in include/*.h is exported by the library that
  implements it;
* host-side pieces restated in the product (initial state, obstacle mask,
  hencol parameters, namelist reader) agree with the oracle / reference data;
* without a GPU the product fails loudly instead of falling back to the CPU.
"""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import oracle_support as O
from bgk3d_gpu_amd import runtime as rt
from bgk3d_gpu_amd.cases import Case, case_from_namelist, read_namelist, sphere_obstacle

REPO = Path(__file__).resolve().parents[1]
GOLD = Path(__file__).parent / "golden" / "reference_data"


def declared(header):
    text = (REPO / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:void|int|char|double|float)\s*\**\s*(\w+)\s*\(", text,
                                 re.M)))


def exported(lib):
    return {name for name in dir(lib)}


@pytest.mark.parametrize("libname,header", [("libbgk3d_gpu_sp.so", "bgk3d_gpu.h"),
                                            ("libbgk3d_gpu_dp.so", "bgk3d_gpu.h"),
                                            ("libbgk3d_engine.so", "bgk3d_engine.h")])
def test_library_exports_every_declared_symbol(libname, header):
    lib = ctypes.CDLL(str(rt.LIBDIR / libname))
    names = declared(header)
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{libname} lacks {missing}"


@pytest.mark.parametrize("libname", ["libbgk3d_gpu_sp.so", "libbgk3d_gpu_dp.so"])
def test_fortran_module_binds_exported_symbols(libname):
    # every interface of the Fortran module names a symbol the C header
    # declares and the library exports, so a Fortran host links against it
    text = (REPO / "bgk3d_gpu_amd" / "fortran" / "bgk3d_gpu_mod.F90").read_text()
    binds = sorted(set(re.findall(r'bind\(\s*C\s*,\s*name\s*=\s*"(\w+)"', text, re.I)))
    assert len(binds) >= 10
    names = set(declared("bgk3d_gpu.h"))
    lib = ctypes.CDLL(str(rt.LIBDIR / libname))
    assert [b for b in binds if b not in names] == []
    assert [b for b in binds if not hasattr(lib, b)] == []
    assert "bgk_init_gpu" in binds and "col_mc_gpu" in binds


def test_reference_symbols_present():
    # exactly the names the reference's Fortran binds (SRC/col_MC.F90:101-123,
    # SRC/bcond_gpu_mod.F90:30-77)
    names = declared("bgk3d_gpu.h")
    for s in ("col_mc_gpu", "bcond_periodic_gpu", "bcond_channel_gpu", "bcond_driven_gpu", "bcond_inflow_gpu",
              "bcond_obs_gpu", "gpu_device_sync"):
        assert s in names


@pytest.mark.parametrize("prec", ["sp", "dp"])
@pytest.mark.parametrize("mode,u00", [(0, 0.1), (1, 0.05)])
def test_host_init_matches_oracle(prec, mode, u00):
    lib = rt.v1(prec)
    R = rt.real_ctype(prec)
    l, m, n = 21, 17, 13
    A = O.new_lattice(l, m, n, prec, ipad=5, jpad=2)
    O.init(A, l, m, n, prec, mode, u00)
    H = O.new_lattice(l, m, n, prec, ipad=5, jpad=2)
    P = (ctypes.c_void_p * 19)(*[H[d].ctypes.data for d in range(19)])
    lib.bgk_init_host.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 6 + [R]
    assert lib.bgk_init_host(P, l, m, n, H.shape[3], H.shape[2], mode, R(u00)) == 0
    assert np.array_equal(A.view(np.uint8), H.view(np.uint8))


@pytest.mark.parametrize("lmn", [(40, 24, 24), (100, 48, 48), (64, 300, 300)])
def test_sphere_obstacle_matches_build_obs(lmn):
    l, m, n = lmn
    obs, box, radius, centre = sphere_obstacle(l, m, n)
    oobs, obox, orad = O.build_obs(l, m, n)
    assert radius == orad and box == obox
    assert np.array_equal(obs, oobs)
    assert centre == (2 * l // 5, m // 2, n // 2)


@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_case_hencol_matches_oracle(prec):
    for svisc, u0, m in ((0.064, 0.1, 64), (0.1, 0.02048, 128), (0.0256, 0.1, 256), (0.01, 0.05, 256)):
        case = Case("channel", 8, m, 8, svisc=svisc, u0=u0)
        om, fg = case.hencol(prec)
        oom, ofg = O.hencol(svisc, u0, m, prec)
        assert om == oom and fg == ofg
        case = Case("ldc", 8, m, 8, svisc=svisc, u0=u0)
        assert case.hencol(prec)[1] == 0.0  # no volume force outside the channel build


def test_namelists_of_reference_tests():
    nml = read_namelist(GOLD / "driven_cavity_256_re1000_bgk.input")
    assert (nml["lx"], nml["ly"], nml["lz"]) == (256, 256, 256)
    assert nml["svisc"] == pytest.approx(0.0256) and nml["u0"] == pytest.approx(0.1)
    assert nml["itfin"] == 100000 and nml["icheck"] == 20000
    case, ctl = case_from_namelist(GOLD / "driven_cavity_256_re1000_bgk.input", "ldc")
    assert case.l == 256 and case.init == 0 and ctl["icheck"] == 20000
    case, _ = case_from_namelist(GOLD / "Sphere_bgk.input", "inflow")
    assert (case.l, case.m, case.n) == (900, 150, 150) and case.obstacle
    for f in GOLD.glob("*_bgk.input"):
        read_namelist(f)


def test_reference_profile_fixture_is_whole():
    blocks = O.parse_blocks((GOLD / "driven_cavity_256_re1000_prof_k.dat").read_text())
    assert sorted(blocks) == [0, 20000, 40000, 60000, 80000, 100000]
    assert all(len(rows) == 256 for rows in blocks.values())
    last = np.array([[float(v) for v in r.split()] for r in blocks[100000]])
    assert last[-1, 1] == pytest.approx(0.0963182, rel=1e-6)


def test_case_validation():
    with pytest.raises(ValueError):
        Case("nope")
    with pytest.raises(ValueError):
        Case("ldc", 0, 4, 4)


@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_reference_abi_refuses_bad_arguments(prec, capfd):
    """Inconsistent sizes or null pointers are refused on the host before any
    launch (the reference would fault on the device): the void entry points
    print and return, the int ones return -1.  No GPU call is made."""
    lib = rt.v1(prec)
    R = rt.real_ctype(prec)
    fake = (ctypes.c_void_p * 19)(*[0x10000 + 0x1000 * d for d in range(19)])
    holed = (ctypes.c_void_p * 19)(*[0x10000 + 0x1000 * d if d != 7 else 0 for d in range(19)])
    out = (ctypes.c_double * 5)()
    l, m, n = 8, 6, 4
    # rows too short for the halo
    assert lib.diagno_gpu(fake, l, m, n, l + 1, m + 2, R(1), out) == -1
    assert "must hold the halo" in capfd.readouterr().err
    # a null population pointer
    assert lib.diagno_gpu(holed, l, m, n, l + 2, m + 2, R(1), out) == -1
    assert "null population pointer" in capfd.readouterr().err
    # non-positive size
    lib.bcond_periodic_gpu(fake, 0, m, n, l + 2, m + 2)
    assert "sizes must be positive" in capfd.readouterr().err
    lib.bcond_driven_gpu(fake, l, m, n, l + 2, m + 1, R(0.01))
    assert "must hold the halo" in capfd.readouterr().err
    lib.col_mc_force_gpu(fake, holed, l, m, n, l + 2, m + 2, R(1.2), R(0), 1)
    assert "null population pointer" in capfd.readouterr().err
    args = [ctypes.c_void_p(0x10000 + 0x1000 * d) for d in range(37)]
    args[20] = ctypes.c_void_p(0)
    lib.col_mc_gpu(*args, l, m, n, l + 2, m + 2, *[R(v) for v in (1.2, 0, 1, 1 / 3, 1 / 18, 1 / 36, 3, 1.5, 3,
                                                                   0, 0, 0)])
    assert "null population pointer" in capfd.readouterr().err
    # obstacle box outside the lattice
    mask = np.zeros((n, m, l), dtype=np.int32)
    lib.bcond_obs_gpu(fake, mask.ctypes.data, l, m, n, l + 2, m + 2, 0, 3, 1, 3, 1, 3, l, m)
    assert "leaves" in capfd.readouterr().err
    u = (ctypes.c_double * n)()
    assert lib.varm_gpu(fake, l, m, n, l + 2, m + 2, R(1), u, u, u, None) == -1
    assert lib.prof_k_gpu(fake, l, m, n, l, m + 2, 2, 2, R(1), (R * (4 * n))()) == -1
    capfd.readouterr()


def test_no_cpu_fallback_without_gpu():
    from conftest import gpu_present
    if gpu_present():
        pytest.skip("a GPU is present")
    from bgk3d_gpu_amd.engine import Engine
    with pytest.raises(rt.BGKError):
        Engine(Case("ldc", 8, 8, 8), "f32")
    assert rt.device_count() == 0
