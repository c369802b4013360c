"""Device-side initial state against the host one (SRC/init.F90:31-209).

bgk_init_gpu (reference-compatible ABI) and the handle API's init form
each cell on the device from per-coordinate sin/cos factors evaluated on the
host; the host form (bgk_init_host) is itself pinned to the oracle's init in
test_abi_host.  The bar is bit-identical arrays, halo included, with the
row and plane padding left as it was.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle_support as O
from bgk3d_gpu_amd import runtime as rt
from bgk3d_gpu_amd.cases import Case
from bgk3d_gpu_amd.engine import Engine
from bgk3d_gpu_amd.simulation import Simulation

pytestmark = pytest.mark.gpu

SHAPES = [(21, 17, 13, 5, 2), (64, 9, 7, 0, 0)]


def banner(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def host_init(prec, l, m, n, ipad, jpad, mode, u00):
    lib = rt.v1(prec)
    R = rt.real_ctype(prec)
    H = O.new_lattice(l, m, n, prec, ipad=ipad, jpad=jpad)
    P = (ctypes.c_void_p * 19)(*[H[d].ctypes.data for d in range(19)])
    lib.bgk_init_host.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 6 + [R]
    assert lib.bgk_init_host(P, l, m, n, H.shape[3], H.shape[2], mode, R(u00)) == 0
    return H


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_reference_abi_device_init_equals_host(gpu, prec, mode, shape):
    l, m, n, ipad, jpad = shape
    u00 = 0.05
    banner(f"bgk_init_gpu {prec} mode={mode} {l}x{m}x{n} ipad={ipad} jpad={jpad}")
    H = host_init(prec, l, m, n, ipad, jpad, mode, u00)
    sim = Simulation(Case("periodic", l, m, n, svisc=0.01, u0=u00, ipad=ipad, jpad=jpad), prec)
    try:
        # overwrite the whole arrays with the padding value first, so that a
        # cell the device init misses (or padding it touches) shows
        sim.set_state(np.full_like(H, np.finfo(H.dtype).max))
        sim.init_device(mode, u00)
        G = sim.state()
    finally:
        sim.close()
    assert np.array_equal(G.view(np.uint8), H.view(np.uint8))


def test_reference_abi_device_init_refuses_bad_mode(gpu):
    sim = Simulation(Case("periodic", 8, 6, 4, svisc=0.01, u0=0.05), "sp")
    try:
        sim.lib.bgk_init_gpu.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 6 + [ctypes.c_float]
        assert sim.lib.bgk_init_gpu(sim.pa, 8, 6, 4, sim.nx, sim.ny, 3, ctypes.c_float(0.05)) == -1
        assert sim.lib.bgk_init_gpu(sim.pa, 8, 6, 4, 9, sim.ny, 1, ctypes.c_float(0.05)) == -1
    finally:
        sim.close()


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("storage", ["f32", "f64", "f16", "bf16"])
def test_handle_device_init_equals_host(gpu, storage, mode):
    l, m, n, u00 = 21, 17, 13, 0.05
    banner(f"handle init {storage} mode={mode} {l}x{m}x{n}")
    prec = "dp" if storage == "f64" else "sp"
    H = host_init(prec, l, m, n, 0, 0, mode, u00)[:, 1:n + 1, 1:m + 1, 1:l + 1]
    with Engine(Case("periodic", l, m, n, svisc=0.01, u0=u00, init=mode), storage) as eng:
        G = eng.populations()
    if storage == "f16":
        H = H.astype(np.float16)
    elif storage == "bf16":
        # round to nearest even on the upper 16 bits of the FP32 pattern
        b = H.astype(np.float32).view(np.uint32).astype(np.uint64)
        b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
        H = b.astype(np.uint32).view(np.float32)
    assert np.array_equal(G, H.astype(np.float64))
