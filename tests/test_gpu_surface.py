"""Drop-in behaviours of the PETSc surface that the reference documents as
its own (its quirks included), checked on the GPU against the oracle or
against what the reference's code does."""

import numpy as np
import pytest

from _helpers import same_bits
from mpk import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def O(oracle):
    oracle.set_mode("dev")
    yield oracle
    oracle.set_mode("ref")


def _mat(csr):
    from mojopetsc4py import PETSc
    ip, ix, va = csr
    n = len(ip) - 1
    A = PETSc.Mat().createSeqAIJ((n, n), nnz=7)
    A.setValuesCSR(ip, ix, va)
    A.assemble()
    return A


def test_vec_sub_returns_minus_other(native):
    """Vec.__sub__ gives -other, not self - other (Vec.py:490-495)."""
    from mojopetsc4py import PETSc
    a = PETSc.Vec().createWithArray(np.arange(5.0) + 1.0)
    b = PETSc.Vec().createWithArray(np.full(5, 0.5))
    assert same_bits((a - b).getArray(), -np.full(5, 0.5))


def test_vec_operators(native):
    """The other operators: -a is a * -1, a + b, a * b and a / b elementwise
    (a / b starts from a zero vector, as the reference's does, and divides
    a by b into it); abs(a) is a zero vector, because the reference takes
    abs() of duplicate() (Vec.py:466-469); a @ b is the dot product."""
    from mojopetsc4py import PETSc
    x = np.random.default_rng(8).normal(size=37)
    y = np.random.default_rng(9).normal(size=37)
    a, b = PETSc.Vec().createWithArray(x), PETSc.Vec().createWithArray(y)
    assert same_bits((-a).getArray(), x * -1.0)
    assert same_bits((+a).getArray(), x)
    assert same_bits((a + b).getArray(), x + y)
    assert same_bits((a * b).getArray(), x * y)
    assert same_bits((a / b).getArray(), x / y)
    assert same_bits(abs(a).getArray(), np.zeros(37))
    assert (a @ b) == a.dot(b)
    assert a.__add__(3.0) is NotImplemented


def test_sor_iterations_are_not_used(native):
    """setSORIterations is stored but the apply runs its = lits = 1
    (PC.py:280-282, capi.mojo:277-280): the same bits as the default."""
    from mojopetsc4py import PETSc
    A = _mat(W.diffusion_3d(10)[0])
    n = A.getSize()[0]
    x = PETSc.Vec().createWithArray(np.random.default_rng(2).normal(size=n))
    out = []
    for its in (None, 5):
        pc = PETSc.PC().create()
        pc.setType("sor")
        if its:
            pc.setSORIterations(its)
        pc.setOperators(A)
        y = PETSc.Vec().createSeq(n)
        pc.apply(x, y)
        out.append(y.getArray().copy())
    assert same_bits(out[0], out[1])


def test_symmetric_jacobi_apply(native, O):
    """applySymmetricLeft / Right with Jacobi multiply by 1/sqrt(|d|) (0 -> 1),
    the square root the set up stores (PC.py:511-514)."""
    from mojopetsc4py import PETSc
    ip, ix, va = W.diffusion_3d(8)[0]
    A = _mat((ip, ix, va))
    n = A.getSize()[0]
    pc = PETSc.PC().create()
    pc.setType("jacobi")
    pc.setOperators(A)
    pc.setUp()
    xh = np.random.default_rng(3).normal(size=n)
    want = xh * O.PCSetup(O.CSR(ip, ix, va), "jacobi").diagsqrt[:n]
    for f in (pc.applySymmetricLeft, pc.applySymmetricRight):
        y = PETSc.Vec().createSeq(n)
        f(PETSc.Vec().createWithArray(xh), y)
        assert same_bits(y.getArray(), want)


def _solve(csr, b, ksp_type, pc_type, **opt):
    from mojopetsc4py import PETSc
    A = _mat(csr)
    n = A.getSize()[0]
    ksp = PETSc.KSP().create()
    ksp.setType(ksp_type)
    ksp.getPC().setType(pc_type)
    ksp.setOperators(A)
    if "side" in opt:
        ksp.setPCSide(opt.pop("side"))
    ksp.setTolerances(**opt)
    x = PETSc.Vec().createSeq(n)
    ksp.solve(PETSc.Vec().createWithArray(b), x)
    return x.getArray().copy(), ksp.getIterationNumber(), ksp.getConvergedReason()


@pytest.mark.parametrize("ksp_type", ["cg", "gmres", "bcgs"])
def test_symmetric_side_with_jacobi(native, O, ksp_type):
    """PC_SYMMETRIC with Jacobi.  PCApplySymmetric reads the square root from
    pci[5] (ksp.mojo:205-206), a slot the reference's set up leaves null
    (PC.py:511-514), so the reference (and the oracle, which restates it)
    dereferences a null pointer.  The product reads pci[2], where the set up
    stores it; the oracle given a descriptor with pci[5] filled the same way
    is the check: its bits."""
    csr, b = W.diffusion_3d(12)
    x, its, reason = _solve(csr, b, ksp_type, "jacobi", side=2, rtol=1e-10, max_it=500)
    A = O.CSR(*csr)
    pc = O.PCSetup(A, "jacobi")
    pc.pci[5] = pc.pci[2]
    r = O.ksp_solve(A, b, ksp_type, pc, side=2, rtol=1e-10, max_it=500)
    assert (its, reason) == (r.its, r.reason)
    assert same_bits(x, r.x)


def test_fgmres_ignores_the_side(native):
    """FGMRES is always right preconditioned with the unpreconditioned norm
    (ksp.mojo:775-868): setting the left side changes no bit."""
    csr, b = W.convection_diffusion_3d(10)
    xl, itl, rl = _solve(csr, b, "fgmres", "ilu", side=0, rtol=1e-10, max_it=500)
    xr, itr, rr = _solve(csr, b, "fgmres", "ilu", side=1, rtol=1e-10, max_it=500)
    assert (itl, rl) == (itr, rr) and same_bits(xl, xr)


def test_factor_matrix_rebuilds_tridiagonal(native):
    """getFactorMatrix returns L U (ILU) or L D L^T (IC) as a matrix
    (PC.py:660-745); on a tridiagonal matrix ILU(0) and IC(0) are exact, so
    the product is A."""
    from mojopetsc4py import PETSc
    ip, ix, va = W.tridiagonal(12)
    A = _mat((ip, ix, va))
    dense = np.zeros((12, 12))
    for i in range(12):
        dense[i, ix[ip[i]:ip[i + 1]]] = va[ip[i]:ip[i + 1]]
    for kind in ("ilu", "icc"):
        pc = PETSc.PC().create()
        pc.setType(kind)
        pc.setOperators(A)
        F = pc.getFactorMatrix()
        fi, fj, fv = F.getValuesCSR()
        got = np.zeros((12, 12))
        for i in range(12):
            got[i, fj[fi[i]:fi[i + 1]]] = fv[fi[i]:fi[i + 1]]
        assert np.allclose(got, dense, rtol=1e-13, atol=1e-13), kind


def test_uncovered_options_raise(native):
    """What the reference does not cover raises NotImplementedError, as it
    does (PC.py:160-168,214-267; KSP.py:168-177)."""
    from mojopetsc4py import PETSc
    pc = PETSc.PC().create()
    with pytest.raises(NotImplementedError):
        pc.setType("spai")
    with pytest.raises(NotImplementedError):
        pc.setFactorOrdering("nd")
    with pytest.raises(NotImplementedError):
        pc.setFactorPivot(1e-3)
    ksp = PETSc.KSP().create()
    with pytest.raises(NotImplementedError):
        ksp.setType("minres")
