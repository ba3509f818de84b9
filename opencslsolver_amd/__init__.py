"""OpenCSL MI355X: the sparse-Krylov + FVM-assembly hot path of OpenCSLSolver
on gfx950 HIP kernels, behind a C ABI (include/ocsl.h)."""
__version__ = "0.1.0"
