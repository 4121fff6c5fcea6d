"""Sequential AIJ matrices, mirroring `petsc4py/PETSc/Mat.pyx` (the
reference's python/mojopetsc4py/PETSc/Mat.py): the same methods, sizes,
defaults and errors.

Host side: CSR arrays (row starts, columns ascending within a row, float64
values) plus a log of the entry writes made since they were last assembled;
assembly folds the log into new arrays with numpy (a stable sort by entry,
then each entry's writes in call order).  `setValuesCSR` installs arrays
directly and keeps int32 index arrays as they are.  Device side: the
products run on the GPU over a copy with int32 row pointers and columns and a
sliced copy (64-row slices) for coalesced loads, uploaded on first use and
again after any change.
"""

from __future__ import annotations

import os

import numpy as np

from mpk import _native
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

    @classmethod
    def adopt(cls, rp, ci, v, nnz, nrows, ncols):
        """A device copy over buffers already on the device (the device fold
        and dense compaction write them)."""
        self = cls.__new__(cls)
        self.rp, self.ci, self.v = rp, ci, v
        self.nnz, self.nrows, self.ncols = int(nnz), nrows, ncols
        self.sellc = None
        self.tsellc = None
        return self

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
            if padded > SELL_MAX_SLOTS:
                self.tsellc = False
                return None
            soff = DeviceBuffer(4 * ((nc + 63) // 64 + 1))
            sv = DeviceBuffer(8 * max(padded, 1))
            sc = DeviceBuffer(6 * max(padded, 1))  # int32 columns, then int16 offsets
            if L.mpk_sell_build(trp_d.ptr, trow_d.ptr, vt_d.ptr, nc, soff.ptr, sv.ptr, sc.ptr) != 0:
                check()
            L.mpk_synchronize()
            self.tsellc = (trp_d, soff, sv, sc)
        return self.tsellc or None


# ---------------------------------------------------------------- host storage

class _HostCSR:
    """The host arrays of an assembled matrix: row starts, columns and values
    with room for `len(values)` entries, of which the first `nnz` are used.
    Index arrays keep the integer width they were handed in (int32 CSR is
    uploaded as it is); `wide()` gives the int64 view the reference exposes."""

    __slots__ = ("indptr", "indices", "values", "nnz", "_wide")

    def __init__(self, indptr, indices, values, nnz):
        self.indptr = indptr
        self.indices = indices
        self.values = values
        self.nnz = int(nnz)
        self._wide = None

    @classmethod
    def empty(cls, nrows: int, room: int = 1):
        room = max(int(room), 1)
        return cls(np.zeros(nrows + 1, np.int64), np.zeros(room, np.int64), np.zeros(room), 0)

    def wide(self):
        """(indptr, indices) as int64 arrays (cached until the arrays change)."""
        if self._wide is None:
            if self.indptr.dtype == np.int64 and self.indices.dtype == np.int64:
                self._wide = (self.indptr, self.indices)
            else:
                self._wide = (self.indptr.astype(np.int64), self.indices.astype(np.int64))
        return self._wide

    def changed(self):
        self._wide = None

    def row_of_entries(self):
        """The row of every used entry, in storage order."""
        ip = self.indptr.astype(np.int64, copy=False)
        return np.repeat(np.arange(len(ip) - 1, dtype=np.int64), np.diff(ip))

    def find(self, row: int, col: int) -> int:
        """Storage position of (row, col), or -1 (a binary search over the
        row's columns, which assembly keeps ascending)."""
        lo, hi = int(self.indptr[row]), int(self.indptr[row + 1])
        k = lo + int(np.searchsorted(self.indices[lo:hi], col))
        return k if k < hi and int(self.indices[k]) == col else -1


class _WriteLog:
    """Entry writes not folded into the CSR arrays yet, in call order: one
    chunk per call of rows, columns, values and whether the call adds.  The
    CSR entries present when the first write arrives count as writes before
    all of them (`base`), so the fold's result is the one a per-entry dict
    gives: an insert replaces, an add accumulates onto the entry's value so
    far (from 0.0 when it has none), and entries keep whatever they last
    became, zeros included."""

    __slots__ = ("rows", "cols", "vals", "adds", "base")

    def __init__(self, base: _HostCSR | None):
        self.rows, self.cols, self.vals, self.adds = [], [], [], []
        self.base = base

    def push(self, rows, cols, vals, add: bool):
        if len(rows):
            self.rows.append(np.asarray(rows, np.int64))
            self.cols.append(np.asarray(cols, np.int64))
            self.vals.append(np.asarray(vals, np.float64))
            self.adds.append(np.full(len(rows), bool(add)))

    def _stream(self):
        """Every write in order: (rows, cols, vals, adds), the base first."""
        r, c, v, a = list(self.rows), list(self.cols), list(self.vals), list(self.adds)
        if self.base is not None and self.base.nnz:
            b = self.base
            ip, ix = b.wide()
            r.insert(0, b.row_of_entries())
            c.insert(0, ix[: b.nnz].copy())
            v.insert(0, b.values[: b.nnz].copy())
            a.insert(0, np.zeros(b.nnz, bool))
        if not r:
            z = np.zeros(0, np.int64)
            return z, z, np.zeros(0), np.zeros(0, bool)
        return np.concatenate(r), np.concatenate(c), np.concatenate(v), np.concatenate(a)

    def value(self, row: int, col: int) -> float:
        """The entry's value as the fold would leave it (0.0 if never set)."""
        acc = 0.0
        b = self.base
        if b is not None and b.nnz and 0 <= row < len(b.indptr) - 1:
            lo, hi = int(b.indptr[row]), int(b.indptr[row + 1])
            hit = np.flatnonzero(b.indices[lo:hi] == col)
            if len(hit):
                acc = float(b.values[lo + hit[-1]])
        for r, c, v, a in zip(self.rows, self.cols, self.vals, self.adds):
            for k in np.flatnonzero((r == row) & (c == col)):
                acc = acc + float(v[k]) if a[k] else float(v[k])
        return acc

    def fold(self, nrows: int, ncols: int) -> _HostCSR:
        """The CSR arrays after every write: rows in order, columns ascending,
        one entry per (row, column) ever written."""
        r, c, v, a = self._stream()
        n = len(r)
        if n == 0:
            return _HostCSR.empty(nrows)
        key = r * max(ncols, 1) + c
        order = np.argsort(key, kind="stable")  # call order kept within a key
        key, v, a = key[order], v[order], a[order]
        r, c = r[order], c[order]
        head = np.ones(n, bool)
        head[1:] = key[1:] != key[:-1]
        starts = np.flatnonzero(head)
        ends = np.append(starts[1:], n)
        out = v[starts].copy()
        # an entry written once takes its value (an add onto 0.0 is the value
        # itself); entries written several times fold their writes in order
        first_add = a[starts]
        out[first_add] = 0.0 + v[starts][first_add]
        multi = np.flatnonzero(ends - starts > 1)
        for g in multi:
            acc = 0.0
            for k in range(starts[g], ends[g]):
                acc = acc + float(v[k]) if a[k] else float(v[k])
            out[g] = acc
        ur, uc = r[starts], c[starts]
        indptr = np.zeros(nrows + 1, np.int64)
        np.cumsum(np.bincount(ur, minlength=nrows), out=indptr[1:])
        m = len(ur)
        indices = np.zeros(max(m, 1), np.int64)
        values = np.zeros(max(m, 1))
        indices[:m] = uc
        values[:m] = out
        return _HostCSR(indptr, indices, values, m)


def _shape_of(size):
    if isinstance(size, (tuple, list)):
        return int(size[0]), int(size[1])
    return int(size), int(size)


def _outer_writes(rows, cols, vals=None, value=None):
    """The (row, col, value) writes of a rows x cols block in row-major order."""
    rows = np.asarray(list(rows), np.int64).ravel()
    cols = np.asarray(list(cols), np.int64).ravel()
    rr = np.repeat(rows, len(cols))
    cc = np.tile(cols, len(rows))
    if vals is None:
        vv = np.full(len(rr), float(value))
    else:
        vv = vals[: len(rr)]
    return rr, cc, vv


class Mat:
    """A sequential AIJ (CSR) matrix: host CSR arrays (assembled on demand
    from the writes made since the last fold) and a device copy for the
    products."""

    Type = MatType
    InfoType = MatInfoType
    OrderingType = MatOrderingType
    Structure = MatStructures
    SORType = MatSORType

    def __init__(self):
        self._m = 0
        self._n = 0
        self._csr: _HostCSR | None = None  # None: not created (or destroyed)
        self._log: _WriteLog | None = None
        self._assembled = False
        self._devc = None
        self._version = 0

    # ---------------------------------------------------------- internal state
    @property
    def _nrows(self):
        return self._m

    @property
    def _ncols(self):
        return self._n

    def _touch(self):
        """The entries changed: the device copy and set ups built on it are stale."""
        self._devc = None
        self._version += 1
        if self._csr is not None:
            self._csr.changed()

    def _materialize(self):
        """Fold the pending writes into the CSR arrays: on the device
        (mpk_mat_fold) for logs of at least MPK_FOLD_DEVICE_MIN writes (65536
        by default) when a device is open, else on the host; the same arrays
        bit for bit either way."""
        if self._log is None:
            return
        folded = self._fold_on_device()
        if folded is None:
            self._csr = self._log.fold(self._m, self._n)
            self._log = None
            self._touch()
            return
        self._csr, devc = folded
        self._log = None
        self._touch()
        self._devc = devc

    def _fold_on_device(self):
        log = self._log
        nw = sum(len(r) for r in log.rows) + (log.base.nnz if log.base is not None else 0)
        if nw < _env_int("MPK_FOLD_DEVICE_MIN", 65536) or nw >= 2 ** 31 or not _native.device_ready():
            return None
        r, c, v, a = log._stream()
        r = np.ascontiguousarray(r, np.int64)
        c = np.ascontiguousarray(c, np.int64)
        v = np.ascontiguousarray(v, np.float64)
        a = np.ascontiguousarray(a, np.uint8)
        rp, ci, vd = DeviceBuffer(4 * (self._m + 1)), DeviceBuffer(4 * max(nw, 1)), DeviceBuffer(8 * max(nw, 1))
        nnz = lib().mpk_mat_fold(r.ctypes.data, c.ctypes.data, v.ctypes.data, a.ctypes.data, nw, self._m, self._n,
                                 rp.ptr, ci.ptr, vd.ptr)
        check()
        return _downloaded(rp, ci, vd, nnz, self._m, self._n)

    def _writes(self) -> _WriteLog:
        if self._log is None:
            self._log = _WriteLog(self._csr)
        return self._log

    def _host(self) -> _HostCSR:
        self._materialize()
        return self._csr

    def _dev(self) -> _DevCSR:
        h = self._host()
        if self._devc is None:
            if self._m:
                cols = h.indices[: h.nnz]
                if len(cols) and (int(cols.min()) < 0 or int(cols.max()) >= self._n):
                    raise ValueError("a column index lies outside the matrix")
            self._devc = _DevCSR(h.indptr, h.indices, h.values, self._m, self._n)
        return self._devc

    def _check(self):
        if self._csr is None:
            raise RuntimeError("Mat has not been created")
        return self

    # ---------------------------------------------------------- creation
    def create(self, comm=None) -> "Mat":
        return self

    def createSeqAIJ(self, size, bsize=None, nnz=None, comm=None) -> "Mat":
        """An m x n matrix with room for `nnz` entries a row and none set."""
        self._m, self._n = _shape_of(size)
        room = self._m * (1 if nnz is None else int(nnz))
        self._csr = _HostCSR.empty(self._m, room)
        self._log = _WriteLog(None)  # nothing written yet: assembling gives no entries
        self._assembled = True
        self._touch()
        return self

    def createAIJ(self, size, bsize=None, nnz=None, comm=None) -> "Mat":
        return self.createSeqAIJ(size, bsize, nnz, comm)

    def setSizes(self, size, bsize=None) -> "Mat":
        self._m, self._n = _shape_of(size)
        if self._csr is None:
            self._csr = _HostCSR.empty(self._m)
            self._log = _WriteLog(None)
        return self

    def duplicate(self, copy: bool = False) -> "Mat":
        h = self._host()
        twin = Mat()
        twin._m, twin._n = self._m, self._n
        twin._assembled = self._assembled
        vals = h.values.copy() if copy else np.zeros_like(h.values)
        twin._csr = _HostCSR(h.indptr.copy(), h.indices.copy(), vals, h.nnz)
        return twin

    def copy(self, result=None, structure=None) -> "Mat":
        """Copy into `result` (a duplicate when None).  With
        DIFFERENT_NONZERO_PATTERN the target keeps its own pattern and takes
        the source's value at each of its entries (0.0 where the source has
        none), as MatCopy does (Mat.py:163-210 of the reference)."""
        src = self._host()
        if result is None:
            result = self.duplicate()
        if structure == MatStructures.DIFFERENT_NONZERO_PATTERN:
            dst = result._host()
            if (result._m, result._n) != (self._m, self._n):
                raise ValueError(f"incompatible sizes: {self.getSize()} into {result.getSize()}")
            width = max(self._n, 1)
            skey = src.row_of_entries() * width + src.wide()[1][: src.nnz]
            order = np.argsort(skey, kind="stable")
            skey, sval = skey[order], src.values[: src.nnz][order]
            last = np.ones(len(skey), bool)  # the last of equal keys wins
            last[:-1] = skey[1:] != skey[:-1]
            skey, sval = skey[last], sval[last]
            dkey = dst.row_of_entries() * width + dst.wide()[1][: dst.nnz]
            at = np.searchsorted(skey, dkey)
            hit = at < len(skey)
            hit[hit] = skey[at[hit]] == dkey[hit]
            fresh = np.zeros(dst.nnz)
            fresh[hit] = sval[at[hit]]
            dst.values[: dst.nnz] = fresh
            result._touch()
            return result
        result._m, result._n = self._m, self._n
        result._csr = _HostCSR(src.indptr.copy(), src.indices.copy(), src.values.copy(), src.nnz)
        result._log = None
        result._assembled = self._assembled
        result._touch()
        return result

    def destroy(self) -> "Mat":
        self._csr = None
        self._log = None
        self._touch()
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
        """Every stored entry becomes 0.0; the pattern stays."""
        h = self._check()._host()
        h.values[:] = 0.0
        self._touch()

    def setPreallocationNNZ(self, nnz) -> "Mat":
        return self

    def setOption(self, option, flag) -> None:
        return None

    def setType(self, mat_type) -> None:
        return None

    def getType(self) -> str:
        return MatType.SEQAIJ

    # ---------------------------------------------------------- accessors
    def getSize(self):
        return (self._m, self._n)

    def getLocalSize(self):
        return self.getSize()

    def getSizes(self):
        return ((self._m, self._m), (self._n, self._n))

    def getOwnershipRange(self):
        return (0, self._m)

    def getOwnershipRangeColumn(self):
        return (0, self._n)

    def getBlockSize(self) -> int:
        return 1

    def getNNZ(self) -> int:
        return self._host().nnz

    def getInfo(self, info=None) -> dict:
        """Counts: the room the arrays hold and the entries in use (stored
        zeros included)."""
        h = self._host()
        return {MatInfoType.NNZ_ALLOCATED: len(h.values), MatInfoType.NZ_USED: h.nnz,
                MatInfoType.ZERO_TOLERANCE: 0.0}

    @property
    def irow(self) -> np.ndarray:
        return self._host().wide()[0]

    @property
    def icol(self) -> np.ndarray:
        return self._host().wide()[1]

    @property
    def data(self) -> np.ndarray:
        return self._host().values

    def getValuesCSR(self):
        h = self._host()
        ip, ix = h.wide()
        return (ip[: self._m + 1].copy(), ix[: h.nnz].copy(), h.values[: h.nnz].copy())

    def setValuesCSR(self, indptr, indices, data=None) -> "Mat":
        """Install CSR arrays (row starts, columns, values).  int32 index
        arrays are kept as they are and uploaded without a wider copy."""
        def ints(a):
            a = np.asarray(a)
            return a if a.dtype in (np.int32, np.int64) else a.astype(np.int64)

        ip, ix = ints(indptr), ints(indices)
        vals = np.zeros(0) if data is None else np.asarray(data, dtype=np.float64)
        if len(ip) != self._m + 1:
            raise ValueError(f"incompatible array sizes: indptr has {len(ip)} entries,"
                             f" {self._m + 1} are needed")
        if len(ix) != len(vals):
            raise ValueError(f"incompatible array sizes: {len(ix)} indices but {len(vals)} values")
        nnz = int(ip[-1])
        old = self._csr
        room = len(old.values) if old is not None and len(old.values) >= nnz else nnz
        if room == nnz:
            cols, values = np.array(ix[:nnz], copy=True), np.array(vals[:nnz], dtype=np.float64)
        else:  # the arrays keep their room past the entries
            cols = np.zeros(room, dtype=ix.dtype)
            values = np.zeros(room)
            cols[:nnz] = ix[:nnz]
            values[:nnz] = vals[:nnz]
        self._csr = _HostCSR(np.array(ip, copy=True), cols, values, nnz)
        self._log = None
        self._touch()
        return self

    def getValue(self, row, col) -> float:
        row, col = int(row), int(col)
        if self._log is not None:
            return self._log.value(row, col)
        h = self._csr
        if h is None or not (0 <= row < self._m and 0 <= col < self._n):
            return 0.0
        k = h.find(row, col)
        return 0.0 if k < 0 else float(h.values[k])

    def getValues(self, rows, cols, values=None) -> np.ndarray:
        rows, cols = list(rows), list(cols)
        block = np.array([[self.getValue(r, c) for c in cols] for r in rows], dtype=np.float64)
        if values is None:
            return block
        values[:] = block
        return values

    def _dense(self) -> np.ndarray:
        h = self._host()
        out = np.zeros((self._m, self._n))
        np.add.at(out, (h.row_of_entries(), h.wide()[1][: h.nnz]), h.values[: h.nnz])
        return out

    def _write(self, rr, cc, vv, add: bool):
        """Record writes in order; the first one out of range raises, after
        the writes before it are kept."""
        bad = np.flatnonzero((rr < 0) | (rr >= self._m) | (cc < 0) | (cc >= self._n))
        upto = int(bad[0]) if len(bad) else len(rr)
        self._writes().push(rr[:upto], cc[:upto], vv[:upto], add)
        self._touch()
        if len(bad):
            raise ValueError(f"entry ({int(rr[upto])}, {int(cc[upto])}) is out of range")

    def setValue(self, row, col, value, addv=None) -> None:
        self._write(np.array([int(row)]), np.array([int(col)]), np.array([float(value)]), addv is not None)

    def setValues(self, rows, cols, values, addv=None, insertmode=None) -> None:
        rows, cols = list(rows), list(cols)
        vals = np.asarray(values, dtype=np.float64).ravel()
        need = len(rows) * len(cols)
        if len(vals) < need:
            raise ValueError(f"incompatible array sizes: ni={len(rows)}, nj={len(cols)},"
                             f" nv={len(vals)}; at least {need} values are needed")
        self._write(*_outer_writes(rows, cols, vals), addv is not None)

    def setValueBlocked(self, rows, cols, value, addv=None) -> None:
        self._write(*_outer_writes(rows, cols, value=value), addv is not None)

    def setValuesBlocked(self, rows, cols, values, addv=None) -> None:
        self.setValues(rows, cols, values, addv)

    # ---------------------------------------------------------- algebra
    def _out(self, x: Vec) -> Vec:
        if x.getSize() != self._n:
            raise ValueError(f"incompatible sizes: matrix {self._n} columns, vector {x.getSize()}")
        return Vec().createSeq(self._m, x.getBlockSize())

    def _outT(self, x: Vec) -> Vec:
        if x.getSize() != self._m:
            raise ValueError(f"incompatible sizes: matrix {self._m} rows, vector {x.getSize()}")
        return Vec().createSeq(self._n, x.getBlockSize())

    def _check_y(self, y: Vec, n: int) -> Vec:
        if y.getSize() != n:
            raise ValueError(f"incompatible sizes: output has {y.getSize()} entries, {n} are needed")
        return y

    def _product(self, x: Vec, out: Vec, add: int):
        d = self._dev()
        s = d.sell()
        L = lib()
        if s is None:
            fn = L.mp_mat_mult_add if add else L.mp_mat_mult
            fn(x._d(), out._dw(), d.rp.ptr, d.ci.ptr, d.v.ptr, self._m)
        else:
            L.mpk_sell_mult(x._d(), out._dw(), d.rp.ptr, s[0].ptr, s[1].ptr, s[2].ptr, self._m, add)
        check()

    def mult(self, x: Vec, y: Vec = None) -> Vec:
        """y = A @ x on the GPU."""
        self._materialize()
        out = self._out(x) if y is None else self._check_y(y, self._m)
        if x.getSize() != self._n:
            raise ValueError(f"incompatible sizes: matrix {self._n} columns, vector {x.getSize()}")
        self._product(x, out, 0)
        return out

    def multAdd(self, x: Vec, v: Vec, y: Vec = None) -> Vec:
        """y = A @ x + v (the products' sums start from v's entries)."""
        self._materialize()
        out = self._out(x) if y is None else self._check_y(y, self._m)
        self._check_y(v, self._m)
        if out is x:
            raise ValueError("output y may not alias the input x in multAdd")
        if out is not v:
            v.copy(out)
        self._product(x, out, 1)
        return out

    def multTranspose(self, x: Vec, y: Vec = None) -> Vec:
        """y = A^T @ x."""
        self._materialize()
        out = self._outT(x) if y is None else self._check_y(y, self._n)
        if x.getSize() != self._m:
            raise ValueError(f"incompatible sizes: matrix {self._m} rows, vector {x.getSize()}")
        if not self._n:
            return out
        if not self._m:
            out.zeroEntries()
            return out
        d = self._dev()
        h = self._csr
        t = d.transpose_sell(h.indptr, h.indices, h.values)
        if t is None:  # past the sliced copy's slots: the transpose-map product
            lib().mp_mat_mult_transpose(x._d(), out._dw(), d.rp.ptr, d.ci.ptr, d.v.ptr, self._m, self._n)
        else:
            # the sliced copy of A^T (MatMultTranspose_SeqAIJ's order, ported.mojo:408-431)
            lib().mpk_sell_mult(x._d(), out._dw(), t[0].ptr, t[1].ptr, t[2].ptr, t[3].ptr, self._n, 0)
        check()
        return out

    def multTransposeAdd(self, x: Vec, v: Vec, y: Vec = None) -> Vec:
        """y = A^T @ x + v."""
        self._materialize()
        out = self._outT(x) if y is None else self._check_y(y, self._n)
        self._check_y(v, self._n)
        self._check_y(x, self._m)
        if out is x:
            raise ValueError("output y may not alias the input x in multTransposeAdd")
        keep = v.copy() if out is v else v  # the product overwrites its output first
        self.multTranspose(x, out)
        out.axpy(1.0, keep)
        return out

    def createVecLeft(self) -> Vec:
        return Vec().createSeq(self._m)

    def createVecRight(self) -> Vec:
        return Vec().createSeq(self._n)

    def createVecs(self):
        return (self.createVecRight(), self.createVecLeft())

    def __call__(self, x, y=None):
        return self.mult(x, self.createVecLeft() if y is None else y)

    def getDiagonal(self, result: Vec = None) -> Vec:
        self._materialize()
        out = Vec().createSeq(self._m) if result is None else self._check_y(result, self._m)
        d = self._dev()
        lib().mp_mat_get_diagonal(out._dw(), d.rp.ptr, d.ci.ptr, d.v.ptr, self._m)
        check()
        return out

    def createDiagonal(self, d) -> "Mat":
        n = d.getSize()
        diag = Mat().createSeqAIJ(n, nnz=1)
        diag.setValuesCSR(np.arange(n + 1), np.arange(n), d.getArray(readonly=True))
        return diag

    def createTranspose(self, mat=None) -> "Mat":
        return _transpose(self if mat is None else mat)

    def transpose(self, out: "Mat" = None) -> "Mat":
        return _transpose(self)

    def scale(self, alpha) -> None:
        """a_ij * alpha for every stored entry."""
        h = self._check()._host()
        h.values *= float(alpha)
        self._touch()

    def equal(self, mat, tol: float = 0.0) -> bool:
        """Same shape, same stored arrays, values within tol."""
        a, b = self._host(), mat._host()
        if self.getSize() != mat.getSize():
            return False
        same = np.array_equal(a.wide()[0], b.wide()[0]) and np.array_equal(a.wide()[1], b.wide()[1])
        return bool(same and np.allclose(a.values, b.values, rtol=0, atol=tol))

    def isSymmetric(self, tol: float = 0.0) -> bool:
        if self._m != self._n:
            return False
        dense = self._dense()
        return bool(np.allclose(dense, dense.T, rtol=0, atol=tol))

    def isTranspose(self, mat=None, tol: float = 0.0) -> bool:
        other = self if mat is None else mat
        if (self._m, self._n) != (other._n, other._m):
            return False
        return bool(np.allclose(self._dense(), other._dense().T, rtol=0, atol=tol))

    def __matmul__(self, x):
        return self.mult(x)

    def __mul__(self, other):
        return _matmul(self, other) if isinstance(other, Mat) else NotImplemented

    def __neg__(self):
        flipped = self.duplicate(copy=True)
        flipped.scale(-1.0)
        return flipped


def _transpose(m: Mat) -> Mat:
    """A^T as CSR, each column's entries in increasing row."""
    h = m._host()
    rows = h.row_of_entries()
    cols = h.wide()[1][: h.nnz]
    order = np.lexsort((rows, cols))
    t = Mat().createSeqAIJ((m._n, m._m), nnz=1)
    indptr = np.zeros(m._n + 1, np.int64)
    if h.nnz:
        np.cumsum(np.bincount(cols, minlength=m._n), out=indptr[1:])
    t.setValuesCSR(indptr, rows[order], h.values[: h.nnz][order])
    return t


def _matmul(a: Mat, b: Mat) -> Mat:
    if a._n != b._m:
        raise ValueError(f"incompatible sizes: {a.getSize()} times {b.getSize()}")
    return from_dense(a._dense() @ b._dense())


def _env_int(name: str, default: int) -> int:
    try:
        return int(os.environ.get(name, default))
    except ValueError:
        return default


def _downloaded(rp, ci, v, nnz, nrows, ncols):
    """(host arrays, device copy) of CSR arrays the device wrote."""
    nnz = int(nnz)
    host = _HostCSR(rp.to_numpy(np.int32, nrows + 1), ci.to_numpy(np.int32, max(nnz, 1)),
                    v.to_numpy(np.float64, max(nnz, 1)), nnz)
    return host, _DevCSR.adopt(rp, ci, v, nnz, nrows, ncols)


def from_dense(array, mat=None) -> Mat:
    """A CSR matrix of the nonzeros of a dense 2-D array: compacted on the
    device (mpk_dense_to_csr) for arrays of at least MPK_DENSE_DEVICE_MIN
    elements (2^20 by default) when a device is open, else on the host."""
    arr = _f64(array)
    if arr.ndim != 2:
        raise ValueError("expected a 2-D array")
    n, m = arr.shape
    out = mat if mat is not None else Mat()
    out.createSeqAIJ((n, m), nnz=1)
    if n * m >= max(_env_int("MPK_DENSE_DEVICE_MIN", 1 << 20), 1) and n * m < 2 ** 31 and _native.device_ready():
        L = lib()
        dense = DeviceBuffer.from_array(np.ascontiguousarray(arr))
        rp = DeviceBuffer(4 * (n + 1))
        nnz = -L.mpk_dense_to_csr(dense.ptr, n, m, rp.ptr, 0, 0, 0)
        check()
        ci, vd = DeviceBuffer(4 * max(nnz, 1)), DeviceBuffer(8 * max(nnz, 1))
        got = L.mpk_dense_to_csr(dense.ptr, n, m, rp.ptr, ci.ptr, vd.ptr, nnz)
        check()
        out._csr, devc = _downloaded(rp, ci, vd, got, n, m)
        out._log = None
        out._touch()
        out._devc = devc
        return out
    rr, cc = np.nonzero(arr)
    indptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rr, minlength=n), out=indptr[1:])
    out.setValuesCSR(indptr, cc.astype(np.int64), arr[rr, cc])
    return out
