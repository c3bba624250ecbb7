"""MI355X-native SIPG Poisson/Helmholtz operator and CG solver.

Drop-in for the matrix-free SIPG path of ``ddg.dg2d.sipg`` (apply, loads,
block-Jacobi and multigrid preconditioned CG), executed by hand-written
gfx950 HIP kernels behind the C ABI in ``include/ddgx.h``.

Submodules: ``mesh`` (Mesh2D and builders), ``reference`` (reference
element), ``sipg`` (the drop-in API), ``operator`` (device handles),
``build`` (compiles libddgx.so).
"""

__all__ = ["mesh", "reference", "sipg", "operator"]
