"""MI355X-native lattice Boltzmann wind tunnel: gfx950 HIP kernels behind a C ABI (include/lbm.h).

The Python modules are host-side helpers and a ctypes binding used by the tests and the benchmark;
the Node.js binding lives in cells_amd/node.
"""
from .lbm import Solver, Solver3D, load_library  # noqa: F401
