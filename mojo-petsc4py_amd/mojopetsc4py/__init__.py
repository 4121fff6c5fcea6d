"""Krylov solvers and preconditioners from petsc4py, running on an MI355X.

`from mojopetsc4py import PETSc` gives the reference's drop-in surface; the
vectors, matrices, preconditioners and Krylov iterations live on the GPU.
"""

from . import PETSc  # noqa: F401

__version__ = "0.1.0"
__upstream__ = "petsc4py 3.25.5"

__all__ = ["PETSc"]
