"""Sequential vectors, mirroring `petsc4py/PETSc/Vec.pyx` (the reference's
python/mojopetsc4py/PETSc/Vec.py), stored on the GPU.

The storage is a device buffer of `n` float64 entries.  `getArray()` hands
back a host mirror, synchronised from the device when the device copy is
newer; the mirror is then treated as possibly modified and is uploaded before
the next device operation on the vector (`getArray(readonly=True)` skips
that).  Every arithmetic method runs as a HIP kernel through libmpk.

Sizes, defaults and error behaviour follow the reference: ValueError on size
mismatches, `norm()` returning a pair for NORM_1_AND_2, `dotNorm2` returning
(x.y, x.x).
"""

from __future__ import annotations

import numpy as np

from mpk._native import check, lib
from mpk.device import DeviceBuffer


def _f64(a, copy=False):
    if copy:
        return np.array(a, dtype=np.float64, order="C", copy=True)
    return np.ascontiguousarray(a, dtype=np.float64)


class VecType:
    """The vector type."""

    SEQ = "seq"
    STANDARD = "seq"
    MPI = "mpi"
    SHARED = "shared"
    NEST = "nest"


class NormType:
    """petscvec.h NormType."""

    NORM_1 = 0
    NORM_2 = 1
    NORM_FROBENIUS = 2
    NORM_INFINITY = 3
    NORM_1_AND_2 = 4
    N_1 = NORM_1
    NORM_MAX = NORM_INFINITY


class Vec:
    """A sequential PETSc vector held in GPU memory."""

    Type = VecType

    def __init__(self):
        self._host = None     # numpy mirror
        self._dev = None      # DeviceBuffer
        self._hnew = False    # host mirror may be newer than the device
        self._dnew = False    # device is newer than the host mirror
        self._bsize = 1
        self._scal = np.zeros(2, dtype=np.float64)

    # ---------------------------------------------------------- creation
    def create(self, comm=None) -> "Vec":
        self._host = None
        self._dev = None
        self._bsize = 1
        return self

    def _adopt(self, arr: np.ndarray, bs: int):
        self._host = arr
        self._dev = None
        self._hnew = True
        self._dnew = False
        self._bsize = bs
        return self

    def createSeq(self, size, bsize=None, comm=None) -> "Vec":
        n = int(size[0] if isinstance(size, (tuple, list)) else size)
        bs = 1 if bsize is None else int(bsize)
        if bs < 1:
            raise ValueError(f"invalid block size {bs}")
        if n % bs:
            raise ValueError(f"global size {n} not divisible by block size {bs}")
        return self._adopt(np.zeros(n, dtype=np.float64), bs)

    def createWithArray(self, array, bsize=None, comm=None) -> "Vec":
        arr = _f64(array, copy=True)
        bs = 1 if bsize is None else int(bsize)
        if len(arr) % bs:
            raise ValueError(
                f"array length {len(arr)} is not a multiple of the block size {bs}")
        return self._adopt(arr, bs)

    def duplicate(self, array=None) -> "Vec":
        n = self.getSize()
        if array is None:
            return Vec()._adopt(np.zeros(n, dtype=np.float64), self._bsize)
        arr = _f64(array, copy=True)
        if len(arr) != n:
            raise ValueError(f"incompatible array size: {len(arr)} != {n}")
        return Vec()._adopt(arr, self._bsize)

    def copy(self, result=None) -> "Vec":
        target = self.duplicate() if result is None else result
        self._peer(target)
        lib().mp_vec_copy(self._d(), target._dw(), self._n())
        return target

    def destroy(self) -> "Vec":
        self._host = None
        self._dev = None
        return self

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    # ---------------------------------------------------------- device plumbing
    def _check(self) -> np.ndarray:
        if self._host is None:
            raise RuntimeError("Vec has not been created")
        return self._host

    def _n(self) -> int:
        return len(self._check())

    def _d(self) -> int:
        """Device address for reading: uploads a newer host mirror first."""
        host = self._check()
        if self._dev is None:
            self._dev = DeviceBuffer(host.nbytes)
            self._hnew = True
        if self._hnew:
            self._dev.upload(host)
            self._hnew = False
        return self._dev.ptr

    def _dw(self) -> int:
        """Device address for an operation that writes the vector."""
        p = self._d()
        self._dnew = True
        return p

    def _sync_host(self) -> None:
        if self._dnew and self._dev is not None:
            self._dev.download(self._host)
            check()
            self._dnew = False

    # ---------------------------------------------------------- accessors
    def getArray(self, readonly: bool = False) -> np.ndarray:
        """The vector's entries as a numpy array (a host mirror)."""
        self._check()
        self._sync_host()
        if not readonly:
            self._hnew = True
        return self._host

    def setArray(self, array) -> None:
        cur = self._check()
        arr = _f64(array, copy=True)
        if len(arr) != len(cur):
            raise ValueError(f"incompatible array size: {len(arr)} != {len(cur)}")
        self._sync_host()
        cur[:] = arr
        self._hnew = True
        self._dnew = False

    def placeArray(self, array) -> None:
        self.setArray(array)

    def getSize(self) -> int:
        return self._n()

    def getLocalSize(self) -> int:
        return self._n()

    def getSizes(self):
        n = self._n()
        return (n, n)

    def getBlockSize(self) -> int:
        return self._bsize

    def getOwnershipRange(self):
        return (0, self._n())

    def getType(self) -> str:
        return "seq"

    def getValue(self, index: int) -> float:
        return float(self.getArray(readonly=True)[index])

    def setValue(self, index: int, value: float) -> None:
        self.getArray()[index] = value

    def getValues(self, indices) -> np.ndarray:
        return self.getArray(readonly=True)[np.asarray(indices, dtype=np.int64)]

    def setValues(self, indices, values, addv=None) -> None:
        idx = np.asarray(indices, dtype=np.int64)
        val = _f64(values)
        arr = self.getArray()
        if addv is None:
            arr[idx] = val
        else:
            arr[idx] += val

    # ---------------------------------------------------------- reductions
    def norm(self, norm_type=None):
        """Compute the vector norm; a pair for NORM_1_AND_2."""
        ntype = NormType.NORM_2 if norm_type is None else int(norm_type)
        z = np.zeros(2, dtype=np.float64)
        lib().mp_vec_norm(self._d(), ntype, z.ctypes.data, self._n())
        check()
        both = ntype == NormType.NORM_1_AND_2
        return (float(z[0]), float(z[1])) if both else float(z[0])

    def normBegin(self, norm_type=None) -> None:
        return None

    def normEnd(self, norm_type=None):
        return self.norm(norm_type)

    def dot(self, other: "Vec") -> float:
        self._peer(other)
        lib().mp_vec_dot(self._d(), other._d(), self._scal.ctypes.data, self._n())
        check()
        return float(self._scal[0])

    def tDot(self, other: "Vec") -> float:
        self._peer(other)
        lib().mp_vec_tdot(self._d(), other._d(), self._scal.ctypes.data, self._n())
        check()
        return float(self._scal[0])

    def dotNorm2(self, other: "Vec"):
        """(x . y, x . x) in one pass."""
        self._peer(other)
        s = np.zeros(2, dtype=np.float64)
        lib().mp_vec_dot_norm2(self._d(), other._d(), s.ctypes.data, s.ctypes.data + 8, self._n())
        check()
        return (float(s[0]), float(s[1]))

    def sum(self) -> float:
        lib().mp_vec_sum(self._d(), self._scal.ctypes.data, self._n())
        check()
        return float(self._scal[0])

    def mean(self) -> float:
        lib().mp_vec_mean(self._d(), self._scal.ctypes.data, self._n())
        check()
        return float(self._scal[0])

    def _minmax(self, fn):
        idx = np.zeros(1, dtype=np.int64)
        val = np.zeros(1, dtype=np.float64)
        fn(self._d(), idx.ctypes.data, val.ctypes.data, self._n())
        check()
        return (int(idx[0]), float(val[0]))

    def min(self):
        return self._minmax(lib().mp_vec_min)

    def max(self):
        return self._minmax(lib().mp_vec_max)

    def normalize(self) -> float:
        lib().mp_vec_normalize(self._dw(), self._scal.ctypes.data, self._n())
        check()
        return float(self._scal[0])

    def maxPointwiseDivide(self, other: "Vec") -> float:
        self._peer(other)
        lib().mp_vec_max_pointwise_divide(self._d(), other._d(), self._scal.ctypes.data, self._n())
        check()
        return float(self._scal[0])

    # ---------------------------------------------------------- elementwise
    def _peer(self, *others: "Vec") -> None:
        n = self._n()
        for other in others:
            if other._n() != n:
                raise ValueError(f"incompatible sizes: {n} and {other._n()}")

    def chop(self, tol) -> None:
        lib().mp_vec_chop(self._dw(), self._n(), float(tol))

    def zeroEntries(self) -> None:
        lib().mp_vec_zero_entries(self._dw(), self._n())

    def set(self, alpha) -> None:
        lib().mp_vec_set(self._dw(), float(alpha), self._n())

    def scale(self, alpha) -> None:
        lib().mp_vec_scale(self._dw(), float(alpha), self._n())

    def shift(self, alpha) -> None:
        lib().mp_vec_shift(self._dw(), float(alpha), self._n())

    def swap(self, other: "Vec") -> None:
        self._peer(other)
        lib().mp_vec_swap(self._dw(), other._dw(), self._n())

    def axpy(self, alpha, other: "Vec") -> None:
        """y = alpha*x + y."""
        self._peer(other)
        x = other._d()
        lib().mp_vec_axpy(self._dw(), float(alpha), x, self._n())

    def aypx(self, alpha, other: "Vec") -> None:
        """y = x + alpha*y."""
        self._peer(other)
        x = other._d()
        lib().mp_vec_aypx(self._dw(), float(alpha), x, self._n())

    def axpby(self, alpha, beta, other: "Vec") -> None:
        """y = alpha*x + beta*y."""
        self._peer(other)
        x = other._d()
        lib().mp_vec_axpby(self._dw(), float(alpha), float(beta), x, self._n())

    def axpbypcz(self, alpha, beta, gamma, other, another) -> None:
        self._peer(other, another)
        x, y = other._d(), another._d()
        lib().mp_vec_axpbypcz(self._dw(), float(alpha), float(beta), float(gamma), x, y, self._n())

    def waxpy(self, alpha, other: "Vec", another: "Vec") -> None:
        """w = alpha*x + y."""
        self._peer(other, another)
        x, y = other._d(), another._d()
        lib().mp_vec_waxpy(self._dw(), float(alpha), x, y, self._n())

    def maxpy(self, alphas, vecs) -> None:
        """y = y + sum(alpha[i] * vecs[i])."""
        alph = _f64(alphas)
        if len(alph) != len(vecs):
            raise ValueError("alphas and vecs must have the same length")
        for v in vecs:
            self._peer(v)
        addrs = np.array([v._d() for v in vecs], dtype=np.int64)
        lib().mp_vec_maxpy(self._dw(), len(alph), alph.ctypes.data, addrs.ctypes.data, self._n())

    def _pointwise(self, fn, other, another):
        self._peer(other, another)
        x, y = other._d(), another._d()
        fn(self._dw(), x, y, self._n())

    def pointwiseMult(self, other: "Vec", another: "Vec") -> None:
        self._pointwise(lib().mp_vec_pointwise_mult, other, another)

    def pointwiseDivide(self, other: "Vec", another: "Vec") -> None:
        self._pointwise(lib().mp_vec_pointwise_divide, other, another)

    def pointwiseMin(self, other: "Vec", another: "Vec") -> None:
        self._pointwise(lib().mp_vec_pointwise_min, other, another)

    def pointwiseMax(self, other: "Vec", another: "Vec") -> None:
        self._pointwise(lib().mp_vec_pointwise_max, other, another)

    def pointwiseMaxAbs(self, other: "Vec", another: "Vec") -> None:
        self._pointwise(lib().mp_vec_pointwise_max_abs, other, another)

    def reciprocal(self) -> None:
        lib().mp_vec_reciprocal(self._dw(), self._n())

    def exp(self) -> None:
        lib().mp_vec_exp(self._dw(), self._n())

    def log(self) -> None:
        lib().mp_vec_log(self._dw(), self._n())

    def sqrtabs(self) -> None:
        lib().mp_vec_sqrtabs(self._dw(), self._n())

    def abs(self) -> None:
        lib().mp_vec_abs(self._dw(), self._n())

    def conjugate(self) -> None:
        lib().mp_vec_conjugate(self._dw(), self._n())

    # ---------------------------------------------------------- operators
    # (the binary ones are attached after the class, from _BINARY)
    def __neg__(self):
        flipped = self.copy()
        flipped.scale(-1.0)  # x * -1.0, the reference's set(-1) and pointwise product
        return flipped

    def __pos__(self):
        return self.copy()

    def __abs__(self):
        # the reference takes abs() of duplicate(), a zero vector
        # (python/mojopetsc4py/PETSc/Vec.py:466-469); kept as it is
        return self.duplicate()

    def __rsub__(self, other):
        return other - self if isinstance(other, Vec) else NotImplemented

    def __matmul__(self, other):
        return self.dot(other)

    def __getitem__(self, i):
        return self.getArray(readonly=True)[i]

    def __setitem__(self, i, v):
        self.getArray()[i] = v


# name -> (what the result starts as, how it is then formed from (a, b)).
# The reference starts __sub__ and __truediv__ from duplicate(), a zero
# vector, so a - b is -b (python/mojopetsc4py/PETSc/Vec.py:490-495); kept.
_BINARY = {
    "__add__": ("copy", lambda out, a, b: out.axpy(1.0, b)),
    "__sub__": ("zeros", lambda out, a, b: out.axpy(-1.0, b)),
    "__mul__": ("copy", lambda out, a, b: out.pointwiseMult(a, b)),
    "__truediv__": ("zeros", lambda out, a, b: out.pointwiseDivide(a, b)),
}


def _binary(start, form):
    def operator(self, other):
        if not isinstance(other, Vec):
            return NotImplemented
        out = self.copy() if start == "copy" else self.duplicate()
        form(out, self, other)
        return out

    return operator


for _name, (_start, _form) in _BINARY.items():
    setattr(Vec, _name, _binary(_start, _form))
Vec.__radd__ = Vec.__add__
Vec.__rmul__ = Vec.__mul__
