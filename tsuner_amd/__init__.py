"""MI355X (gfx950) implementation of the TSUNER per-step solver path.

The compute path is libtsuner.so (hand-written HIP kernels behind a C ABI,
include/tsuner.h); this package holds its Python binding (``Solver``), the
build script and the seeded workload generators.
"""
from . import scenarios  # noqa: F401

__all__ = ["scenarios", "Solver", "run_tsuner"]


def __getattr__(name):
    if name in ("Solver", "run_tsuner"):
        from . import solver
        return getattr(solver, name)
    raise AttributeError(name)
