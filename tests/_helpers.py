"""Shared helpers for the parity tests: device upload/download through
libmpk and oracle-side equivalents of the C-ABI calls."""

from __future__ import annotations

import numpy as np


def dev(arr):
    """Upload to a new device buffer.  Keep the returned object alive for as
    long as its address is in use: the buffer is freed with it."""
    from mpk.device import DeviceBuffer
    return DeviceBuffer.from_array(np.ascontiguousarray(arr))


def back(buf, dtype, n):
    return buf.to_numpy(dtype, n)


def bits(x):
    return np.ascontiguousarray(x, dtype=np.float64).view(np.int64)


def same_bits(a, b):
    return np.array_equal(bits(a), bits(b))


def rng_vec(n, seed):
    return np.random.default_rng(seed).normal(size=n)


class DevMat:
    """A matrix uploaded the way the PETSc layer uploads it (int32 CSR)."""

    def __init__(self, indptr, indices, data):
        self.n = len(indptr) - 1
        nnz = int(indptr[-1])
        self.rp = dev(np.asarray(indptr, dtype=np.int32))
        self.ci = dev(np.asarray(indices[:max(nnz, 1)], dtype=np.int32) if nnz else np.zeros(1, np.int32))
        self.v = dev(np.asarray(data[:max(nnz, 1)], dtype=np.float64) if nnz else np.zeros(1))


def wide_band(n=40000, far=35000):
    """SPD, diagonally dominant: 4 on the diagonal, -1 at distance 1 and -0.5
    at distance `far` (so entries lie up to `far` columns from their row:
    past 32767, where sliced copies keep int32 columns)."""
    from mpk import workloads as W
    rows, cols, vals = [], [], []
    for d, v in ((0, 4.0), (1, -1.0), (-1, -1.0), (far, -0.5), (-far, -0.5)):
        i = np.arange(max(0, -d), min(n, n - d))
        rows.append(i)
        cols.append(i + d)
        vals.append(np.full(len(i), v))
    return W._csr_from_coo(n, np.concatenate(rows).astype(np.int64), np.concatenate(cols).astype(np.int64),
                           np.concatenate(vals))
