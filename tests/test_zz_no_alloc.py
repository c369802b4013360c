"""diagno_gpu_noalloc (include/bgk3d_gpu.h): the allocation-free variant of
diagno_gpu, whose partial sums live in a static device array, against
diagno_gpu itself through the C-ABI: two calls in a row, the second on a
smaller lattice (fewer partials than the first left in the array), in FP32
and FP64.  Same kernels and reduction order, so the results are bit-identical.
"""
import sys

import numpy as np
import pytest

from bgk3d_gpu_amd.cases import Case
from bgk3d_gpu_amd.simulation import Simulation

pytestmark = pytest.mark.gpu


def both(sim):
    c = sim.case
    ref = np.zeros(5)
    got = np.full(5, np.nan)
    args = (sim.pa, c.l, c.m, c.n, sim.nx, sim.ny, sim.Rc(1.0))
    assert sim.lib.diagno_gpu(*args, ref.ctypes.data) == 0
    assert sim.lib.diagno_gpu_noalloc(*args, got.ctypes.data) == 0
    return ref, got


@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_noalloc_diagno_equals_diagno(gpu, prec):
    big = Case("ldc", 70, 66, 40, svisc=0.02, u0=0.1)       # 2,640 rows: 1,024 partials
    small = Case("ldc", 12, 10, 8, svisc=0.064, u0=0.1)     # 80 rows: 80 partials
    for case, steps in ((big, 30), (small, 15)):
        print(f"[gpu] diagno_gpu_noalloc {prec} {case.l}x{case.m}x{case.n} steps={steps}", flush=True)
        sys.stdout.flush()
        sim = Simulation(case, prec)
        try:
            sim.step(steps)
            ref, got = both(sim)
        finally:
            sim.close()
        assert np.all(np.isfinite(ref)) and abs(ref[0]) > 0
        assert np.array_equal(got, ref), (got, ref)


def test_noalloc_diagno_rejects_bad_arguments(gpu):
    sim = Simulation(Case("ldc", 8, 8, 8), "sp")
    try:
        out = np.zeros(5)
        c = sim.case
        assert sim.lib.diagno_gpu_noalloc(sim.pa, c.l, c.m, c.n, c.l + 1, sim.ny, sim.Rc(1.0), out.ctypes.data) == -1
        assert sim.lib.diagno_gpu_noalloc(sim.pa, c.l, c.m, c.n, sim.nx, sim.ny, sim.Rc(1.0), None) == -1
    finally:
        sim.close()
