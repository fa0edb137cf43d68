"""p4pdes_amd -- MI355X (gfx950) CG + geometric-multigrid Poisson solver.

The product is a C-ABI shared library (include/p4g.h, built as
p4pdes_amd/lib/libp4g.so) with hand-written HIP kernels, a PETSc-shaped
host surface for DMDA-local callbacks (include/p4g_petsc.h), and the fish
driver (p4pdes_amd/fish).  This package holds the build script and a
ctypes binding used by the tests and the benchmark.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
