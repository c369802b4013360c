"""gfx950-native D3Q19 BGK time step (lattice-Boltzmann), drop-in for the
reference's col_MC / bcond_* / diagno family.

Layers:
  csrc/           hand-written HIP kernels and the C-ABI (include/bgk3d_gpu.h)
  runtime.py      ctypes bindings and device-memory plumbing
  cases.py        run configurations (reference flags + namelist)
  simulation.py   host driver over the reference-compatible ABI
"""
from .cases import Case, case_from_namelist, read_namelist, sphere_obstacle  # noqa: F401

__all__ = ["Case", "case_from_namelist", "read_namelist", "sphere_obstacle"]
