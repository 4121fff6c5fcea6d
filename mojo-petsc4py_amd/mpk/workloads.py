"""Synthetic systems for the parity tests and the benchmark.

Every generator returns int32 CSR (`indptr`, `indices`, `data`) with columns
ascending within each row, plus whatever right-hand side the workload names.
All randomness is seeded here, so a workload is fixed by its name and sizes.

* `laplacian_2d(k)` -- the 5-point Laplacian on a k x k grid, Dirichlet
  boundary eliminated, 4 on the diagonal and -1 off it, row-major numbering,
  each row listing (up, left, self, right, down) -- the matrix of the
  reference's bench (bench/bench.py:64-93).  k = 64 is the 4096-row case.
* `tridiagonal(n)` -- the README usage example's 1-D matrix (4, -1).
* `diffusion_3d(N)` -- 7-point variable-coefficient diffusion on N^3 cells,
  face coefficients lognormal(0, 1) drawn from seed 0 in the order x-faces,
  y-faces, z-faces (boundary faces included, Dirichlet), a_pp = sum of the
  six face coefficients, a_pq = -k_f.  SPD by diagonal dominance.
  Right-hand side uniform[-1, 1] from seed 1.
* `convection_diffusion_3d(N)` -- 7-point -Laplace (6, -1) plus first-order
  upwind convection with velocity (1, 0.5, 0.25) at cell Peclet number
  P_x = v_x h / nu = 10, so P = (10, 5, 2.5): a_pp = 6 + sum(P), the upwind
  neighbour (the -x, -y, -z side, velocity being positive) is -1 - P_d, the
  downwind neighbour -1.  Nonsymmetric.  Right-hand side uniform[-1, 1],
  seed 2.
* `random_spd(n)` -- for each row 12 distinct offsets uniform in
  1 <= j - i <= 5000 (clipped to n), values uniform[-1, 0], mirrored;
  a_ii = 1.05 * row abs-sum.  Seed 3.
* `batch_systems(count)` -- the batched family: sizes log-uniform in
  [200, 20000] (seed 5), rectangles nx = round(sqrt(n rho)), rho log-uniform in
  [1/4, 4], ny = round(n / nx), 2-D 5-point diffusion with lognormal(0, 1) face
  coefficients.
"""

from __future__ import annotations

import numpy as np


def _csr_from_coo(n, rows, cols, vals, ncols=None):
    """Sort (row, col) pairs, sum duplicates and return int32 CSR."""
    ncols = n if ncols is None else ncols
    key = rows.astype(np.int64) * ncols + cols.astype(np.int64)
    order = np.argsort(key, kind="stable")
    key = key[order]
    vals = vals[order]
    uniq, start = np.unique(key, return_index=True)
    summed = np.add.reduceat(vals, start) if len(vals) else vals
    r = (uniq // ncols).astype(np.int64)
    c = (uniq % ncols).astype(np.int32)
    counts = np.bincount(r, minlength=n)
    indptr = np.zeros(n + 1, dtype=np.int32)
    indptr[1:] = np.cumsum(counts)
    return indptr, c, summed.astype(np.float64)


def laplacian_2d(k: int):
    n = k * k
    p = np.arange(n)
    cands = np.stack([
        np.where(p >= k, p - k, -1),
        np.where(p % k > 0, p - 1, -1),
        p,
        np.where(p % k < k - 1, p + 1, -1),
        np.where(p < n - k, p + k, -1),
    ], axis=1).astype(np.int64)
    keep = cands >= 0
    counts = keep.sum(axis=1)
    indptr = np.zeros(n + 1, dtype=np.int32)
    indptr[1:] = np.cumsum(counts)
    indices = cands[keep].astype(np.int32)
    row_of = np.repeat(np.arange(n), counts)
    vals = np.full(indices.shape, -1.0)
    vals[indices == row_of] = 4.0
    return indptr, indices, vals


def tridiagonal(n: int, diag: float = 4.0, off: float = -1.0):
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in sorted({max(i - 1, 0), i, min(i + 1, n - 1)}):
            rows.append(i)
            cols.append(j)
            vals.append(diag if j == i else off)
    return _csr_from_coo(n, np.array(rows), np.array(cols), np.array(vals))


def _stencil7(N, diag, xm, xp, ym, yp, zm, zp):
    """Assemble a 7-point operator on N^3 cells (x fastest) from per-cell
    diagonal and neighbour coefficient arrays shaped (N, N, N) as [z, y, x].
    Neighbours off the grid are dropped.  Columns come out ascending:
    -z, -y, -x, self, +x, +y, +z."""
    n = N ** 3
    idx = np.arange(n, dtype=np.int64).reshape(N, N, N)
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    cols = np.stack([
        np.where(z > 0, idx - N * N, -1),
        np.where(y > 0, idx - N, -1),
        np.where(x > 0, idx - 1, -1),
        idx,
        np.where(x < N - 1, idx + 1, -1),
        np.where(y < N - 1, idx + N, -1),
        np.where(z < N - 1, idx + N * N, -1),
    ], axis=-1).reshape(n, 7)
    vals = np.stack([zm, ym, xm, diag, xp, yp, zp], axis=-1).reshape(n, 7)
    keep = cols >= 0
    counts = keep.sum(axis=1)
    indptr = np.zeros(n + 1, dtype=np.int32)
    indptr[1:] = np.cumsum(counts)
    return indptr, cols[keep].astype(np.int32), vals[keep].astype(np.float64)


def diffusion_3d(N: int, seed: int = 0, rhs_seed: int = 1):
    rng = np.random.default_rng(seed)
    kx = np.exp(rng.standard_normal((N, N, N + 1)))  # [z, y, xface]
    ky = np.exp(rng.standard_normal((N, N + 1, N)))  # [z, yface, x]
    kz = np.exp(rng.standard_normal((N + 1, N, N)))  # [zface, y, x]
    xm, xp = kx[:, :, :-1], kx[:, :, 1:]
    ym, yp = ky[:, :-1, :], ky[:, 1:, :]
    zm, zp = kz[:-1, :, :], kz[1:, :, :]
    diag = ((xm + xp) + (ym + yp)) + (zm + zp)
    A = _stencil7(N, diag, -xm, -xp, -ym, -yp, -zm, -zp)
    b = np.random.default_rng(rhs_seed).uniform(-1.0, 1.0, N ** 3)
    return A, b


def convection_diffusion_3d(N: int, peclet=(10.0, 5.0, 2.5), rhs_seed: int = 2):
    px, py, pz = peclet
    shape = (N, N, N)
    one = np.ones(shape)
    diag = (6.0 + px + py + pz) * one
    A = _stencil7(N, diag, (-1.0 - px) * one, -one, (-1.0 - py) * one, -one,
                  (-1.0 - pz) * one, -one)
    b = np.random.default_rng(rhs_seed).uniform(-1.0, 1.0, N ** 3)
    return A, b


def random_spd(n: int, per_row: int = 12, band: int = 5000, seed: int = 3):
    rng = np.random.default_rng(seed)
    # 12 distinct offsets per row: draw, and redraw the rows holding a repeat
    offs = np.sort(rng.integers(1, band + 1, size=(n, per_row)), axis=1)
    while True:
        bad = np.nonzero((np.diff(offs, axis=1) == 0).any(axis=1))[0]
        if len(bad) == 0:
            break
        offs[bad] = np.sort(rng.integers(1, band + 1, size=(len(bad), per_row)), axis=1)
    r = np.repeat(np.arange(n, dtype=np.int64), per_row)
    c = r + offs.ravel()
    ok = c < n  # clipped to n
    r, c = r[ok], c[ok]
    v = -rng.uniform(0.0, 1.0, size=len(r))
    rr = np.concatenate([r, c])
    cc = np.concatenate([c, r])
    vv = np.concatenate([v, v])
    absrow = np.bincount(rr, weights=np.abs(vv), minlength=n)
    diag = 1.05 * absrow
    diag[diag == 0.0] = 1.0
    rr = np.concatenate([rr, np.arange(n)])
    cc = np.concatenate([cc, np.arange(n)])
    vv = np.concatenate([vv, diag])
    A = _csr_from_coo(n, rr, cc, vv)
    b = np.random.default_rng(seed + 100).uniform(-1.0, 1.0, n)
    return A, b


def diffusion_2d(nx: int, ny: int, rng: np.random.Generator):
    """2-D 5-point diffusion on nx x ny cells (x fastest), lognormal faces."""
    kx = np.exp(rng.standard_normal((ny, nx + 1)))
    ky = np.exp(rng.standard_normal((ny + 1, nx)))
    xm, xp = kx[:, :-1], kx[:, 1:]
    ym, yp = ky[:-1, :], ky[1:, :]
    diag = (xm + xp) + (ym + yp)
    n = nx * ny
    idx = np.arange(n, dtype=np.int64).reshape(ny, nx)
    y, x = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    cols = np.stack([
        np.where(y > 0, idx - nx, -1),
        np.where(x > 0, idx - 1, -1),
        idx,
        np.where(x < nx - 1, idx + 1, -1),
        np.where(y < ny - 1, idx + nx, -1),
    ], axis=-1).reshape(n, 5)
    vals = np.stack([-ym, -xm, diag, -xp, -yp], axis=-1).reshape(n, 5)
    keep = cols >= 0
    counts = keep.sum(axis=1)
    indptr = np.zeros(n + 1, dtype=np.int32)
    indptr[1:] = np.cumsum(counts)
    return indptr, cols[keep].astype(np.int32), vals[keep].astype(np.float64)


def batch_systems(count: int = 4096, seed: int = 5, nmin: int = 200, nmax: int = 20000):
    """The batched family: list of (indptr, indices, data, b, solver) tuples;
    even indices are cg + jacobi, odd ones gmres(20) + ilu(0)."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), count))).astype(np.int64)
    rho = np.exp(rng.uniform(np.log(0.25), np.log(4.0), count))
    out = []
    for s in range(count):
        n = int(sizes[s])
        nx = max(1, int(round(np.sqrt(n * rho[s]))))
        ny = max(1, int(round(n / nx)))
        A = diffusion_2d(nx, ny, rng)
        b = rng.uniform(-1.0, 1.0, nx * ny)
        out.append((A[0], A[1], A[2], b, "cg+jacobi" if s % 2 == 0 else "gmres+ilu"))
    return out
