"""System information, mirroring `petsc4py/PETSc/Sys.pyx` (the reference's
python/mojopetsc4py/PETSc/Sys.py)."""

from __future__ import annotations

import platform
import sys

PETSC_VERSION = (3, 25, 5)
PETSC_VERSION_MAJOR = PETSC_VERSION[0]
PETSC_VERSION_MINOR = PETSC_VERSION[1]
PETSC_VERSION_SUBMINOR = PETSC_VERSION[2]
PETSC_VERSION_STRING = ".".join(str(v) for v in PETSC_VERSION)

Scalar = float
Real = float
Int = int
Bool = bool

__version__ = PETSC_VERSION_STRING


def getVersion() -> tuple[int, int, int]:
    return PETSC_VERSION


def getVersionInfo() -> str:
    return (f"PETSc {PETSC_VERSION_STRING} interface (mojo-petsc4py on MI355X) on"
            f" Python {sys.version.split()[0]} {platform.machine()}")


def getMachine() -> str:
    return platform.machine()


def hasThreads() -> bool:
    return False
