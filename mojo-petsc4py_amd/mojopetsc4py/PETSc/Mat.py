"""Sequential AIJ matrices, mirroring `petsc4py/PETSc/Mat.pyx` (the
reference's python/mojopetsc4py/PETSc/Mat.py).

Assembly stays on the host exactly as in the reference: entries set one by
one collect in a pending dict and become CSR (int64 row starts, int64
columns ascending within a row, float64 values) on demand; `setValuesCSR`
installs CSR directly.  The products run on the GPU over a device copy of the
matrix with int32 row pointers and columns, uploaded on first use and again
after any change to the host arrays.
"""

from __future__ import annotations

import numpy as np

from mpk._native import check, lib
from mpk.device import DeviceBuffer

from .Vec import Vec, _f64

# the most padded slots a sliced copy can hold (int32 slot indices)
SELL_MAX_SLOTS = 2**31 - 1


class MatType:
    """The matrix type."""

    AIJ = "aij"
    SEQAIJ = "seqaij"
    MPIAIJ = "mpiaij"
    SBAIJ = "sbaij"
    DENSE = "dense"
    SHELL = "shell"
    NEST = "nest"


class MatInfoType:
    NNZ = "nz_allocated"
    NZ_USED = "nz_used"
    NNZ_USED = "nz_used"
    NNZ_ALLOCATED = "nz_allocated"
    FACTOR_NONZERO = "factor_nonz"
    ZERO_TOLERANCE = "rcond"
    ROW_ZERO_TOLERANCE = "rcond"


class MatStructures:
    SAME_NONZERO_PATTERN = 2
    DIFFERENT_NONZERO_PATTERN = 0
    SAME = SAME_NONZERO_PATTERN
    DIFFERENT = DIFFERENT_NONZERO_PATTERN


class MatOrderingType:
    NATURAL = "natural"
    ND = "nd"
    RCM = "rcm"
    MD = "md"
    ROWLENGTH = "rowlength"
    QMD = "qmd"
    AMD = "amd"
    METISND = "metisnd"
    EXTERNAL = "external"
    NATURAL_OR_ND = "natural_or_nd"


class MatSORType:
    ZERO_INITIAL_GUESS = 16
    FORWARD_SWEEP = 1
    BACKWARD_SWEEP = 2
    LOCAL_FORWARD_SWEEP = 4
    LOCAL_BACKWARD_SWEEP = 8
    LOCAL_SYMMETRIC_SWEEP = 12
    EISENSTAT = 32
    APPLY_UPPER = 64
    APPLY_LOWER = 128
    FORWARD = 1
    BACKWARD = 2
    SYMMETRIC = 3
    LOCAL_FORWARD = 4
    LOCAL_BACKWARD = 8
    LOCAL_SYMMETRIC = 12


class _DevCSR:
    """The device copy: int32 row pointers and columns, float64 values, and
    (built on first use) the sliced copies of A and of A^T the products read."""

    __slots__ = ("rp", "ci", "v", "nnz", "nrows", "ncols", "sellc", "tsellc")

    def __init__(self, irow, icol, data, nrows, ncols):
        nnz = int(irow[nrows]) if nrows else 0
        if nnz >= 2 ** 31:
            raise ValueError("the device matrix holds int32 indices: nnz must stay below 2**31")
        self.nnz = nnz
        self.nrows = nrows
        self.ncols = ncols
        self.rp = DeviceBuffer.from_array(np.asarray(irow[: nrows + 1], dtype=np.int32))
        self.ci = DeviceBuffer.from_array(np.asarray(icol[: max(nnz, 1)], dtype=np.int32))
        self.v = DeviceBuffer.from_array(np.asarray(data[: max(nnz, 1)], dtype=np.float64))
        self.sellc = None
        self.tsellc = None

    def sell(self):
        """(soff, sv, sc): rows in 64-row slices, column-major inside a slice,
        so a wave's loads are contiguous; row sums stay in CSR order.  None
        when the copy would pass the int32 slot indices (2^31 - 1 padded
        slots): the products then run the CSR row kernel."""
        if self.sellc is None:
            L = lib()
            padded = L.mpk_sell_size(self.rp.ptr, self.nrows)
            check()
            if padded > SELL_MAX_SLOTS:
                self.sellc = False
                return None
            soff = DeviceBuffer(4 * ((self.nrows + 63) // 64 + 1))
            sv = DeviceBuffer(8 * max(padded, 1))
            sc = DeviceBuffer(6 * max(padded, 1))  # int32 columns, then int16 offsets
            if L.mpk_sell_build(self.rp.ptr, self.ci.ptr, self.v.ptr, self.nrows, soff.ptr, sv.ptr, sc.ptr) != 0:
                check()
            self.sellc = (soff, sv, sc)
        return self.sellc or None

    def transpose_sell(self, irow, icol, data):
        """(trp, soff, sv, sc): the sliced copy of A^T, built from the host CSR
        the device copy was made from: column c's entries in increasing row,
        the order MatMultTranspose_SeqAIJ accumulates them in, so a sliced
        product of it gives the transpose product bit for bit."""
        if self.tsellc is None:
            L = lib()
            nr, nc, nnz = self.nrows, self.ncols, self.nnz
            ix = np.asarray(icol[:nnz], dtype=np.int64)
            rows = np.repeat(np.arange(nr, dtype=np.int64), np.diff(np.asarray(irow[: nr + 1], dtype=np.int64)))
            order = np.argsort(ix, kind="stable")
            trp = np.zeros(nc + 1, dtype=np.int32)
            trp[1:] = np.cumsum(np.bincount(ix, minlength=nc))
            trp_d = DeviceBuffer.from_array(trp)
            trow_d = DeviceBuffer.from_array(rows[order].astype(np.int32) if nnz else np.zeros(1, np.int32))
            vt_d = DeviceBuffer.from_array(np.asarray(data[:nnz], dtype=np.float64)[order] if nnz
                                           else np.zeros(1))
            padded = L.mpk_sell_size(trp_d.ptr, nc)
            check()
            soff = DeviceBuffer(4 * ((nc + 63) // 64 + 1))
            sv = DeviceBuffer(8 * max(padded, 1))
            sc = DeviceBuffer(6 * max(padded, 1))  # int32 columns, then int16 offsets
            if L.mpk_sell_build(trp_d.ptr, trow_d.ptr, vt_d.ptr, nc, soff.ptr, sv.ptr, sc.ptr) != 0:
                check()
            L.mpk_synchronize()
            self.tsellc = (trp_d, soff, sv, sc)
        return self.tsellc


class Mat:
    """A sequential AIJ (CSR) matrix."""

    Type = MatType
    InfoType = MatInfoType
    OrderingType = MatOrderingType
    Structure = MatStructures
    SORType = MatSORType

    def __init__(self):
        self._irow = None
        self._icol = None
        self._data = None
        self._nrows = 0
        self._ncols = 0
        self._assembled = False
        self._nnz_used = 0
        self._pending = None
        self._devc = None
        self._version = 0

    # ---------------------------------------------------------- device copy
    def _touch(self):
        """The host arrays changed: the device copy and any set up built on
        it are stale."""
        self._devc = None
        self._version += 1

    def _dev(self) -> _DevCSR:
        self._materialize()
        if self._devc is None:
            if self._nrows:
                cols = self._icol[: int(self._irow[self._nrows])]
                if len(cols) and (cols.min() < 0 or cols.max() >= self._ncols):
                    raise ValueError("a column index lies outside the matrix")
            self._devc = _DevCSR(self._irow, self._icol, self._data, self._nrows, self._ncols)
        return self._devc

    # ---------------------------------------------------------- creation
    def create(self, comm=None) -> "Mat":
        return self

    def createSeqAIJ(self, size, bsize=None, nnz=None, comm=None) -> "Mat":
        if isinstance(size, (tuple, list)):
            m, n = int(size[0]), int(size[1])
        else:
            m = n = int(size)
        per_row = 1 if nnz is None else int(nnz)
        self._nrows = m
        self._ncols = n
        self._irow = np.arange(m + 1, dtype=np.int64) * per_row
        self._icol = np.zeros(m * per_row, dtype=np.int64)
        self._data = np.zeros(m * per_row, dtype=np.float64)
        self._assembled = True
        self._pending = {}
        self._touch()
        return self

    def createAIJ(self, size, bsize=None, nnz=None, comm=None) -> "Mat":
        return self.createSeqAIJ(size, bsize, nnz, comm)

    def setSizes(self, size, bsize=None) -> "Mat":
        if isinstance(size, (tuple, list)):
            self._nrows, self._ncols = int(size[0]), int(size[1])
        else:
            self._nrows = self._ncols = int(size)
        if self._irow is None:
            self._irow = np.zeros(self._nrows + 1, dtype=np.int64)
            self._icol = np.zeros(1, dtype=np.int64)
            self._data = np.zeros(1, dtype=np.float64)
            self._pending = {}
        return self

    def duplicate(self, copy: bool = False) -> "Mat":
        self._materialize()
        out = Mat()
        out._nrows = self._nrows
        out._ncols = self._ncols
        out._assembled = self._assembled
        out._nnz_used = self._nnz_used
        out._irow = self._irow.copy()
        out._icol = self._icol.copy()
        out._data = self._data.copy() if copy else np.zeros_like(self._data)
        return out

    def copy(self, result=None, structure=None) -> "Mat":
        """Copy into `result`; with DIFFERENT_NONZERO_PATTERN the target keeps
        its pattern and takes the values of the entries it holds
        (Mat.py:163-210 of the reference)."""
        source = self
        source._materialize()
        target = source.duplicate() if result is None else result
        if structure == MatStructures.DIFFERENT_NONZERO_PATTERN:
            target._materialize()
            if (target._nrows, target._ncols) != (source._nrows, source._ncols):
                raise ValueError(
                    f"incompatible sizes: {source.getSize()} into {target.getSize()}")
            src = {
                (int(source._icol[j]), i): float(source._data[j])
                for i in range(source._nrows)
                for j in range(int(source._irow[i]), int(source._irow[i + 1]))
            }
            trow, tcol, _ = target.getValuesCSR()
            newdata = np.zeros(len(tcol), dtype=np.float64)
            for i in range(target._nrows):
                for j in range(int(trow[i]), int(trow[i + 1])):
                    newdata[j] = src.get((int(tcol[j]), i), 0.0)
            target._data[: len(newdata)] = newdata
            target._touch()
            return target
        target._nrows = source._nrows
        target._ncols = source._ncols
        target._irow = source._irow.copy()
        target._icol = source._icol.copy()
        target._data = source._data.copy()
        target._nnz_used = source._nnz_used
        target._assembled = source._assembled
        target._pending = None
        target._touch()
        return target

    def destroy(self) -> "Mat":
        self._irow = None
        self._touch()
        return self

    def _check(self):
        if self._irow is None:
            raise RuntimeError("Mat has not been created")
        return self

    # ---------------------------------------------------------- assembly
    def setUp(self) -> "Mat":
        self._check()._materialize()
        self._assembled = True
        return self

    def assemblyBegin(self, assembly=None) -> None:
        return None

    def assemblyEnd(self, assembly=None) -> None:
        self._assembled = True

    def assemble(self, assembly=None) -> None:
        self.assemblyEnd()

    def zeroEntries(self) -> None:
        self._check()
        self._materialize()
        self._data[:] = 0.0
        self._touch()

    def setPreallocationNNZ(self, nnz) -> "Mat":
        return self

    def setOption(self, option, flag) -> None:
        return None

    def setType(self, mat_type) -> None:
        return None

    def getType(self) -> str:
        return "seqaij"

    # ---------------------------------------------------------- accessors
    def getSize(self):
        return (self._nrows, self._ncols)

    def getLocalSize(self):
        return (self._nrows, self._ncols)

    def getSizes(self):
        return ((self._nrows, self._nrows), (self._ncols, self._ncols))

    def getOwnershipRange(self):
        return (0, self._nrows)

    def getOwnershipRangeColumn(self):
        return (0, self._ncols)

    def getBlockSize(self) -> int:
        return 1

    def getNNZ(self) -> int:
        self._materialize()
        return int(self._nnz_used)

    def getInfo(self, info=None) -> dict:
        self._materialize()
        return {"nz_allocated": int(len(self._data)), "nz_used": int(self._nnz_used), "rcond": 0.0}

    @property
    def irow(self) -> np.ndarray:
        self._materialize()
        return self._irow

    @property
    def icol(self) -> np.ndarray:
        self._materialize()
        return self._icol

    @property
    def data(self) -> np.ndarray:
        self._materialize()
        return self._data

    def getValuesCSR(self):
        self._materialize()
        nnz = int(self._nnz_used)
        return (self._irow[: self._nrows + 1].copy(), self._icol[:nnz].copy(), self._data[:nnz].copy())

    def setValuesCSR(self, indptr, indices, data=None) -> "Mat":
        """Install CSR arrays (row starts, columns, values) directly."""
        irow = np.asarray(indptr, dtype=np.int64)
        icol = np.asarray(indices, dtype=np.int64)
        vals = np.zeros(0, dtype=np.float64) if data is None else np.asarray(data, dtype=np.float64)
        if len(irow) != self._nrows + 1:
            raise ValueError(
                f"incompatible array sizes: indptr has {len(irow)} entries,"
                f" {self._nrows + 1} are needed")
        if len(icol) != len(vals):
            raise ValueError(f"incompatible array sizes: {len(icol)} indices but {len(vals)} values")
        nnz = int(irow[-1])
        if self._data is None or self._data.size < nnz:
            self._data = np.zeros(nnz, dtype=np.float64)
            self._icol = np.zeros(nnz, dtype=np.int64)
        if self._irow is None or len(self._irow) != self._nrows + 1:
            self._irow = np.zeros(self._nrows + 1, dtype=np.int64)
        self._irow[:] = irow
        self._icol[:nnz] = icol
        self._data[:nnz] = vals
        self._nnz_used = nnz
        self._pending = None
        self._touch()
        return self

    def _locate(self, row: int, col: int):
        if row < 0 or row >= self._nrows or col < 0 or col >= self._ncols:
            return (-1, -1)
        lo = int(self._irow[row])
        hi = int(self._irow[row + 1])
        pos = int(np.searchsorted(self._icol[lo:hi], col))
        if pos >= hi - lo or int(self._icol[lo + pos]) != col:
            return (-1, -1)
        return (row, lo + pos)

    def getValue(self, row, col) -> float:
        return self._entry(int(row), int(col))

    def getValues(self, rows, cols, values=None) -> np.ndarray:
        rows = list(rows)
        cols = list(cols)
        out = np.array([[self.getValue(r, c) for c in cols] for r in rows], dtype=np.float64)
        if values is not None:
            values[:] = out
            return values
        return out

    def _dense(self) -> np.ndarray:
        self._materialize()
        out = np.zeros((self._nrows, self._ncols), dtype=np.float64)
        for i in range(self._nrows):
            for j in range(int(self._irow[i]), int(self._irow[i + 1])):
                out[i, int(self._icol[j])] += float(self._data[j])
        return out

    def _materialize(self):
        """Pending entries -> CSR, columns sorted within each row."""
        if self._pending is None:
            return
        rows = self._pending
        nrow = self._nrows
        counts = np.zeros(nrow, dtype=np.int64)
        for r, cols in rows.items():
            counts[r] = len(cols)
        irow = np.zeros(nrow + 1, dtype=np.int64)
        irow[1:] = np.cumsum(counts)
        total = int(irow[-1])
        icol = np.zeros(max(total, 1), dtype=np.int64)
        data = np.zeros(max(total, 1), dtype=np.float64)
        for r, cols in rows.items():
            k = int(irow[r])
            for c in sorted(cols):
                icol[k] = c
                data[k] = cols[c]
                k += 1
        self._irow = irow
        self._icol = icol
        self._data = data
        self._nnz_used = total
        self._pending = None
        self._touch()

    def _entry(self, row: int, col: int) -> float:
        if self._pending is not None:
            return self._pending.get(row, {}).get(col, 0.0)
        i, j = self._locate(row, col)
        return 0.0 if i < 0 else float(self._data[j])

    def _spill(self):
        if self._pending is not None:
            return
        self._pending = {}
        for i in range(self._nrows):
            cols = {}
            for j in range(int(self._irow[i]), int(self._irow[i + 1])):
                cols[int(self._icol[j])] = float(self._data[j])
            if cols:
                self._pending[i] = cols

    def _setValue(self, row: int, col: int, value: float, addv: bool):
        if row < 0 or row >= self._nrows or col < 0 or col >= self._ncols:
            raise ValueError(f"entry ({row}, {col}) is out of range")
        self._spill()
        cols = self._pending.setdefault(row, {})
        cols[col] = cols.get(col, 0.0) + value if addv else value
        self._touch()

    def setValue(self, row, col, value, addv=None) -> None:
        self._setValue(int(row), int(col), float(value), addv is not None)

    def setValues(self, rows, cols, values, addv=None, insertmode=None) -> None:
        rows = list(rows)
        cols = list(cols)
        vals = list(np.asarray(values, dtype=np.float64).ravel())
        need = len(rows) * len(cols)
        if len(vals) < need:
            raise ValueError(
                f"incompatible array sizes: ni={len(rows)}, nj={len(cols)},"
                f" nv={len(vals)}; at least {need} values are needed")
        add = addv is not None
        for i, r in enumerate(rows):
            for j, c in enumerate(cols):
                self._setValue(int(r), int(c), float(vals[i * len(cols) + j]), add)

    def setValueBlocked(self, rows, cols, value, addv=None) -> None:
        add = addv is not None
        for r in list(rows):
            for c in list(cols):
                self._setValue(int(r), int(c), float(value), add)

    def setValuesBlocked(self, rows, cols, values, addv=None) -> None:
        self.setValues(rows, cols, values, addv)

    # ---------------------------------------------------------- algebra
    def _out(self, x: Vec) -> Vec:
        if x.getSize() != self._ncols:
            raise ValueError(
                f"incompatible sizes: matrix {self._ncols} columns, vector {x.getSize()}")
        return Vec().createSeq(self._nrows, x.getBlockSize())

    def _outT(self, x: Vec) -> Vec:
        if x.getSize() != self._nrows:
            raise ValueError(
                f"incompatible sizes: matrix {self._nrows} rows, vector {x.getSize()}")
        return Vec().createSeq(self._ncols, x.getBlockSize())

    def _check_y(self, y: Vec, n: int) -> Vec:
        if y.getSize() != n:
            raise ValueError(f"incompatible sizes: output has {y.getSize()} entries, {n} are needed")
        return y

    def mult(self, x: Vec, y: Vec = None) -> Vec:
        """y = A @ x on the GPU."""
        self._materialize()
        out = self._out(x) if y is None else self._check_y(y, self._nrows)
        if x.getSize() != self._ncols:
            raise ValueError(f"incompatible sizes: matrix {self._ncols} columns, vector {x.getSize()}")
        d = self._dev()
        xa = x._d()
        s = d.sell()
        if s is None:
            lib().mp_mat_mult(xa, out._dw(), d.rp.ptr, d.ci.ptr, d.v.ptr, self._nrows)
        else:
            lib().mpk_sell_mult(xa, out._dw(), d.rp.ptr, s[0].ptr, s[1].ptr, s[2].ptr, self._nrows, 0)
        check()
        return out

    def multAdd(self, x: Vec, v: Vec, y: Vec = None) -> Vec:
        """y = A @ x + v."""
        self._materialize()
        out = self._out(x) if y is None else self._check_y(y, self._nrows)
        self._check_y(v, self._nrows)
        if out is x:
            raise ValueError("output y may not alias the input x in multAdd")
        if out is not v:
            v.copy(out)
        d = self._dev()
        xa = x._d()
        s = d.sell()
        if s is None:
            lib().mp_mat_mult_add(xa, out._dw(), d.rp.ptr, d.ci.ptr, d.v.ptr, self._nrows)
        else:
            lib().mpk_sell_mult(xa, out._dw(), d.rp.ptr, s[0].ptr, s[1].ptr, s[2].ptr, self._nrows, 1)
        check()
        return out

    def multTranspose(self, x: Vec, y: Vec = None) -> Vec:
        """y = A^T @ x."""
        self._materialize()
        out = self._outT(x) if y is None else self._check_y(y, self._ncols)
        if x.getSize() != self._nrows:
            raise ValueError(f"incompatible sizes: matrix {self._nrows} rows, vector {x.getSize()}")
        d = self._dev()
        xa = x._d()
        if self._ncols:
            if self._nrows:
                # the sliced copy of A^T (MatMultTranspose_SeqAIJ's order, ported.mojo:408-431)
                trp, so, sv, sc = d.transpose_sell(self._irow, self._icol, self._data)
                lib().mpk_sell_mult(xa, out._dw(), trp.ptr, so.ptr, sv.ptr, sc.ptr, self._ncols, 0)
            else:
                out.zeroEntries()
        check()
        return out

    def multTransposeAdd(self, x: Vec, v: Vec, y: Vec = None) -> Vec:
        """y = A^T @ x + v."""
        self._materialize()
        out = self._outT(x) if y is None else self._check_y(y, self._ncols)
        self._check_y(v, self._ncols)
        self._check_y(x, self._nrows)
        if out is x:
            raise ValueError("output y may not alias the input x in multTransposeAdd")
        if out is v:
            v = v.copy()
        self.multTranspose(x, out)
        out.axpy(1.0, v)
        return out

    def createVecLeft(self) -> Vec:
        return Vec().createSeq(self._nrows)

    def createVecRight(self) -> Vec:
        return Vec().createSeq(self._ncols)

    def createVecs(self):
        return (self.createVecRight(), self.createVecLeft())

    def __call__(self, x, y=None):
        if y is None:
            y = self.createVecLeft()
        self.mult(x, y)
        return y

    def getDiagonal(self, result: Vec = None) -> Vec:
        self._materialize()
        out = Vec().createSeq(self._nrows) if result is None else self._check_y(result, self._nrows)
        d = self._dev()
        lib().mp_mat_get_diagonal(out._dw(), d.rp.ptr, d.ci.ptr, d.v.ptr, self._nrows)
        check()
        return out

    def createDiagonal(self, d) -> "Mat":
        n = d.getSize()
        out = Mat().createSeqAIJ(n, nnz=1)
        rows = np.arange(n, dtype=np.int64)
        out.setValuesCSR(np.arange(n + 1), rows, d.getArray(readonly=True))
        return out

    def createTranspose(self, mat=None) -> "Mat":
        return _transpose(self if mat is None else mat)

    def transpose(self, out: "Mat" = None) -> "Mat":
        return _transpose(self)

    def scale(self, alpha) -> None:
        """Scale every stored entry (the same product the device kernel
        forms, a_ij * alpha)."""
        self._check()
        self._materialize()
        self._data *= float(alpha)
        self._touch()

    def equal(self, mat, tol: float = 0.0) -> bool:
        other = mat
        self._materialize()
        other._materialize()
        if self._nrows != other._nrows or self._ncols != other._ncols:
            return False
        if not np.array_equal(self._irow, other._irow):
            return False
        if not np.array_equal(self._icol, other._icol):
            return False
        return bool(np.allclose(self._data, other._data, rtol=0, atol=tol))

    def isSymmetric(self, tol: float = 0.0) -> bool:
        if self._nrows != self._ncols:
            return False
        return bool(np.allclose(self._dense(), _transpose(self)._dense(), rtol=0, atol=tol))

    def isTranspose(self, mat=None, tol: float = 0.0) -> bool:
        other = self if mat is None else mat
        if self._nrows != other._ncols or self._ncols != other._nrows:
            return False
        return bool(np.allclose(self._dense(), _transpose(other)._dense(), rtol=0, atol=tol))

    def __matmul__(self, x):
        return self.mult(x)

    def __mul__(self, other):
        if not isinstance(other, Mat):
            return NotImplemented
        return _matmul(self, other)

    def __neg__(self):
        out = self.duplicate(copy=True)
        out.scale(-1.0)
        return out


def _transpose(m: Mat) -> Mat:
    """Transpose a CSR matrix (columns of the result ascending)."""
    rows, cols, vals = m.getValuesCSR()
    out = Mat().createSeqAIJ(m._ncols, nnz=1)
    out._ncols = m._nrows
    counts = np.bincount(cols.astype(np.int64), minlength=m._ncols) if len(cols) else np.zeros(m._ncols, np.int64)
    irow = np.zeros(m._ncols + 1, dtype=np.int64)
    irow[1:] = np.cumsum(counts)
    row_of = np.repeat(np.arange(m._nrows, dtype=np.int64), np.diff(rows))
    order = np.lexsort((row_of, cols))
    out._irow = irow
    out._icol = row_of[order] if len(order) else np.zeros(1, dtype=np.int64)
    out._data = vals[order] if len(order) else np.zeros(1, dtype=np.float64)
    out._nnz_used = int(irow[-1])
    out._assembled = True
    out._pending = None
    out._touch()
    return out


def _matmul(a: Mat, b: Mat) -> Mat:
    if a._ncols != b._nrows:
        raise ValueError(f"incompatible sizes: {a.getSize()} times {b.getSize()}")
    return from_dense(a._dense() @ b._dense())


def from_dense(array, mat=None) -> Mat:
    """Build a CSR matrix from a dense 2-D array (nonzeros only)."""
    arr = _f64(array)
    if arr.ndim != 2:
        raise ValueError("expected a 2-D array")
    n, m = arr.shape
    out = mat if mat is not None else Mat()
    out.createSeqAIJ((n, m), nnz=1)
    rr, cc = np.nonzero(arr)
    counts = np.count_nonzero(arr, axis=1).astype(np.int64)
    out._irow = np.zeros(n + 1, dtype=np.int64)
    out._irow[1:] = np.cumsum(counts)
    total = int(out._irow[-1])
    out._icol = cc.astype(np.int64) if total else np.zeros(1, dtype=np.int64)
    out._data = arr[rr, cc].astype(np.float64) if total else np.zeros(1, dtype=np.float64)
    out._nnz_used = total
    out._assembled = True
    out._pending = None
    out._touch()
    return out
