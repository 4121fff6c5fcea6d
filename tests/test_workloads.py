"""The synthetic systems: valid CSR, the sizes the configurations name, and
the structural properties each one promises."""

import numpy as np
import pytest

from mpk import workloads as W


def check_csr(ip, ix, va, n):
    assert len(ip) == n + 1 and ip[0] == 0 and np.all(np.diff(ip) >= 0)
    assert len(ix) == ip[-1] == len(va)
    for i in range(0, n, max(1, n // 200)):
        row = ix[ip[i]:ip[i + 1]]
        assert np.all(np.diff(row) > 0) and i in row


def test_laplacian_64_is_config1():
    ip, ix, va = W.laplacian_2d(64)
    check_csr(ip, ix, va, 4096)
    assert ip[-1] == 20224  # 5 * 4096 - 4 * 64
    assert ip.dtype == np.int32 and ix.dtype == np.int32


def test_diffusion_sizes_and_symmetry():
    (ip, ix, va), b = W.diffusion_3d(8)
    n = 512
    check_csr(ip, ix, va, n)
    assert ip[-1] == 7 * n - 6 * 64
    import scipy.sparse as sp
    A = sp.csr_matrix((va, ix, ip), shape=(n, n))
    assert abs(A - A.T).max() == 0.0
    d = A.diagonal()
    off = np.asarray(abs(A).sum(axis=1)).ravel() - d
    assert np.all(d > off - 1e-12)  # diagonally dominant
    assert b.min() >= -1.0 and b.max() <= 1.0


def test_diffusion_128_nnz_formula():
    N = 128
    assert 7 * N ** 3 - 6 * N ** 2 == 14581760


def test_convection_diffusion_upwind():
    (ip, ix, va), _ = W.convection_diffusion_3d(5)
    import scipy.sparse as sp
    A = sp.csr_matrix((va, ix, ip), shape=(125, 125)).toarray()
    c = 2 + 5 * 2 + 25 * 2
    assert A[c, c] == 6 + 10 + 5 + 2.5
    assert A[c, c - 1] == -11.0 and A[c, c + 1] == -1.0
    assert A[c, c - 5] == -6.0 and A[c, c - 25] == -3.5


def test_random_spd_structure():
    (ip, ix, va), b = W.random_spd(3000, per_row=12, band=500)
    n = 3000
    check_csr(ip, ix, va, n)
    import scipy.sparse as sp
    A = sp.csr_matrix((va, ix, ip), shape=(n, n))
    assert abs(A - A.T).max() == 0.0
    offs = np.abs(A.indices - np.repeat(np.arange(n), np.diff(A.indptr)))
    assert offs.max() <= 500
    d = A.diagonal()
    np.testing.assert_allclose(d, 1.05 * (np.asarray(abs(A).sum(axis=1)).ravel() - d), rtol=1e-12)


def test_batch_family():
    sys_ = W.batch_systems(count=64)
    assert len(sys_) == 64
    for s, (ip, ix, va, b, kind) in enumerate(sys_):
        n = len(ip) - 1
        assert 100 <= n <= 25000 and len(b) == n
        assert kind == ("cg+jacobi" if s % 2 == 0 else "gmres+ilu")
