"""SpMV, the Jacobi / SOR set up and apply, and the ILU / IC sweeps on the
GPU against the oracle, bit for bit, through the C ABI."""

import numpy as np
import pytest

from _helpers import DevMat, back, dev, rng_vec, same_bits, wide_band
from mpk import workloads as W

pytestmark = pytest.mark.gpu


def _random_sparse(n, per_row, seed, diag=True):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(n):
        cs = set(rng.integers(0, n, per_row).tolist())
        if diag:
            cs.add(i)
        for c in sorted(cs):
            rows.append(i)
            cols.append(c)
    vals = rng.normal(size=len(rows))
    vals[np.array(rows) == np.array(cols)] += 8.0
    return W._csr_from_coo(n, np.array(rows), np.array(cols), vals)


MATS = {
    "lap64": lambda: W.laplacian_2d(64),
    "diff12": lambda: W.diffusion_3d(12)[0],
    "cd10": lambda: W.convection_diffusion_3d(10)[0],
    "rand500": lambda: _random_sparse(500, 9, 3),
    "tri33": lambda: W.tridiagonal(33),
}


@pytest.fixture(scope="module")
def O(oracle):
    oracle.set_mode("dev")
    yield oracle
    oracle.set_mode("ref")


def _ragged(n, seed):
    """Rows of very uneven length: empty rows, a dense row, short rows."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(n):
        k = 0 if i % 7 == 3 else (n if i == n // 2 else int(rng.integers(1, 12)))
        for c in sorted(set(rng.integers(0, n, k).tolist())):
            rows.append(i)
            cols.append(c)
    return W._csr_from_coo(n, np.array(rows, dtype=np.int64), np.array(cols, dtype=np.int64),
                           rng.normal(size=len(rows)))


SELL_MATS = dict(MATS)
SELL_MATS.update({f"ragged{n}": (lambda n=n: _ragged(n, n)) for n in (1, 63, 64, 65, 130, 1000)})
SELL_MATS["wide40000"] = wide_band  # entries 35000 columns from their row: int32 columns


@pytest.mark.parametrize("nt", ["default", "1"])
@pytest.mark.parametrize("name", sorted(SELL_MATS))
def test_sliced_spmv_bitwise(native, O, monkeypatch, name, nt):
    """The sliced (64-row, column-major) copy gives the CSR row sums bit for
    bit, for y = A x and y = y + A x, including empty rows and partial slices;
    also in the non-temporal form (the default from 3,000,000 rows, forced
    here by MPK_SELL_NT=1).  The copy holds the int32 columns, and 16-bit
    offsets from the row when every entry is within 32767 columns of its row
    (marked by -1 - padded in the last slice offset); the products read the
    offsets then, and the columns otherwise (wide40000)."""
    if nt != "default":
        monkeypatch.setenv("MPK_SELL_NT", nt)
    ip, ix, va = SELL_MATS[name]()
    A = O.CSR(ip, ix, va)
    D = DevMat(ip, ix, va)
    n = A.n
    padded = native.mpk_sell_size(D.rp.ptr, n)
    widths = [max((ip[i + 1] - ip[i]) for i in range(s, min(s + 64, n))) for s in range(0, n, 64)]
    assert padded == 64 * sum(widths)
    so = dev(np.zeros((n + 63) // 64 + 1, np.int32))
    sv = dev(np.zeros(max(padded, 1)))
    sc = dev(np.zeros(3 * max(padded, 1) // 2 + 1, np.int32))  # 6 bytes a slot
    assert native.mpk_sell_build(D.rp.ptr, D.ci.ptr, D.v.ptr, n, so.ptr, sv.ptr, sc.ptr) == 0
    span = max((abs(int(c) - i) for i in range(n) for c in ix[ip[i]:ip[i + 1]]), default=0)
    offs = back(so, np.int32, (n + 63) // 64 + 1)
    assert offs[-1] == (-1 - padded if span <= 32767 else padded)
    raw = back(sc, np.int32, 3 * max(padded, 1) // 2 + 1)
    cols, d16 = raw[:padded], raw.view(np.int16)[2 * padded:3 * padded]
    for i in range(0, n, max(1, n // 97)):  # the slots of a spread of rows
        for k in range(ip[i + 1] - ip[i]):
            q = int(offs[i // 64]) + i % 64 + 64 * k
            assert cols[q] == ix[ip[i] + k]
            if span <= 32767:
                assert int(d16[q]) == int(ix[ip[i] + k]) - i
    x = rng_vec(n, 1)
    y0 = rng_vec(n, 2)
    dx, dy = dev(x), dev(y0)
    native.mpk_sell_mult(dx.ptr, dy.ptr, D.rp.ptr, so.ptr, sv.ptr, sc.ptr, n, 0)
    assert same_bits(back(dy, np.float64, n), A.mult(x))
    dy = dev(y0)
    native.mpk_sell_mult(dx.ptr, dy.ptr, D.rp.ptr, so.ptr, sv.ptr, sc.ptr, n, 1)
    h = y0.copy()
    O.lib().or_mat_mult_add(x.ctypes.data, h.ctypes.data, A.irow.ctypes.data, A.icol.ctypes.data,
                            A.data.ctypes.data, n)
    assert same_bits(back(dy, np.float64, n), h)


def test_sliced_copy_int32_columns_on_request(native, O, monkeypatch):
    """MPK_SELL_COLS=32 at build keeps int32 columns only (the form the bench
    compares with): the last slice offset holds the slot count, and the row
    sums are the same bits."""
    ip, ix, va = W.diffusion_3d(12)[0]
    A = O.CSR(ip, ix, va)
    D = DevMat(ip, ix, va)
    n = A.n
    padded = native.mpk_sell_size(D.rp.ptr, n)
    x = rng_vec(n, 1)
    ys = []
    for cols in ("32", None):
        if cols:
            monkeypatch.setenv("MPK_SELL_COLS", cols)
        else:
            monkeypatch.delenv("MPK_SELL_COLS", raising=False)
        so = dev(np.zeros((n + 63) // 64 + 1, np.int32))
        sv = dev(np.zeros(max(padded, 1)))
        sc = dev(np.zeros(3 * max(padded, 1) // 2 + 1, np.int32))
        assert native.mpk_sell_build(D.rp.ptr, D.ci.ptr, D.v.ptr, n, so.ptr, sv.ptr, sc.ptr) == 0
        assert back(so, np.int32, (n + 63) // 64 + 1)[-1] == (padded if cols else -1 - padded)
        dx, dy = dev(x), dev(np.zeros(n))
        native.mpk_sell_mult(dx.ptr, dy.ptr, D.rp.ptr, so.ptr, sv.ptr, sc.ptr, n, 0)
        ys.append(back(dy, np.float64, n))
    assert same_bits(ys[0], A.mult(x)) and same_bits(ys[1], A.mult(x))


@pytest.mark.parametrize("name", sorted(MATS))
def test_spmv_family_bitwise(native, O, name):
    ip, ix, va = MATS[name]()
    A = O.CSR(ip, ix, va)
    D = DevMat(ip, ix, va)
    n = A.n
    x = rng_vec(n, 1)
    y0 = rng_vec(n, 2)
    dx, dy = dev(x), dev(y0)
    native.mp_mat_mult(dx.ptr, dy.ptr, D.rp.ptr, D.ci.ptr, D.v.ptr, n)
    assert same_bits(back(dy, np.float64, n), A.mult(x))
    dy = dev(y0)
    native.mp_mat_mult_add(dx.ptr, dy.ptr, D.rp.ptr, D.ci.ptr, D.v.ptr, n)
    h = y0.copy()
    O.lib().or_mat_mult_add(x.ctypes.data, h.ctypes.data, A.irow.ctypes.data, A.icol.ctypes.data,
                            A.data.ctypes.data, n)
    assert same_bits(back(dy, np.float64, n), h)
    native.mp_mat_mult_transpose(dx.ptr, dy.ptr, D.rp.ptr, D.ci.ptr, D.v.ptr, n, n)
    O.lib().or_mat_mult_transpose(x.ctypes.data, h.ctypes.data, A.irow.ctypes.data, A.icol.ctypes.data,
                                  A.data.ctypes.data, n, n)
    assert same_bits(back(dy, np.float64, n), h)
    native.mp_mat_get_diagonal(dy.ptr, D.rp.ptr, D.ci.ptr, D.v.ptr, n)
    O.lib().or_mat_get_diagonal(h.ctypes.data, A.irow.ctypes.data, A.icol.ctypes.data, A.data.ctypes.data, n)
    assert same_bits(back(dy, np.float64, n), h)
    dm = dev(np.zeros(n, np.int32))
    dense = native.mp_mat_get_diagonal_markers(D.rp.ptr, D.ci.ptr, dm.ptr, n)
    hm = np.zeros(n, np.int64)
    odense = O.lib().or_mat_get_diagonal_markers(A.irow.ctypes.data, A.icol.ctypes.data, hm.ctypes.data, n)
    assert dense == odense and np.array_equal(back(dm, np.int32, n), hm)


@pytest.mark.parametrize("name", ["lap64", "rand500", "diff12"])
@pytest.mark.parametrize("jtype", [0, 1, 2, 3])
def test_jacobi_setup_bitwise(native, O, name, jtype):
    ip, ix, va = MATS[name]()
    A = O.CSR(ip, ix, va)
    D = DevMat(ip, ix, va)
    n = A.n
    for useabs, fixdiag, scale in ((0, 1, 1.0), (1, 1, 0.5), (0, 0, 2.0)):
        dd, dx = dev(np.zeros(n)), dev(np.zeros(n))
        native.mp_pc_jacobi_setup(dd.ptr, jtype, useabs, fixdiag, scale, D.rp.ptr, D.ci.ptr, D.v.ptr, n, dx.ptr)
        hd, hx = np.zeros(n), np.zeros(n)
        O.lib().or_pc_jacobi_setup(hd.ctypes.data, jtype, useabs, fixdiag, scale, A.irow.ctypes.data,
                                   A.icol.ctypes.data, A.data.ctypes.data, n, hx.ctypes.data)
        assert same_bits(back(dd, np.float64, n), hd)
    native.mp_pc_jacobi_setup_symmetric(dd.ptr, jtype, D.rp.ptr, D.ci.ptr, D.v.ptr, n)
    O.lib().or_pc_jacobi_setup_symmetric(hd.ctypes.data, jtype, A.irow.ctypes.data, A.icol.ctypes.data,
                                         A.data.ctypes.data, n)
    assert same_bits(back(dd, np.float64, n), hd)


@pytest.mark.parametrize("name", ["lap64", "diff12", "cd10", "rand500"])
@pytest.mark.parametrize("omega,fshift,flag", [(1.0, 0.0, 12 | 16), (1.2, 0.0, 12 | 16), (0.8, 0.1, 1 | 16),
                                              (1.0, 0.0, 2 | 16)])
def test_sor_apply_bitwise(native, O, name, omega, fshift, flag):
    ip, ix, va = MATS[name]()
    A = O.CSR(ip, ix, va)
    D = DevMat(ip, ix, va)
    n = A.n
    b = rng_vec(n, 4)
    hm = np.zeros(n, np.int64)
    O.lib().or_mat_get_diagonal_markers(A.irow.ctypes.data, A.icol.ctypes.data, hm.ctypes.data, n)
    dm = dev(hm.astype(np.int32))
    dy, di, dd, dt = dev(np.zeros(n)), dev(np.zeros(n)), dev(np.zeros(n)), dev(np.zeros(3 * n))
    db = dev(b)  # keep the buffer alive across the call
    native.mp_pc_sor_apply(db.ptr, dy.ptr, D.rp.ptr, D.ci.ptr, D.v.ptr, dm.ptr, di.ptr, dd.ptr, dt.ptr, omega,
                           fshift, flag, n)
    hy, hi, hd, ht = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(3 * n)
    O.lib().or_pc_sor_apply(b.ctypes.data, hy.ctypes.data, A.irow.ctypes.data, A.icol.ctypes.data,
                            A.data.ctypes.data, hm.ctypes.data, hi.ctypes.data, hd.ctypes.data, ht.ctypes.data,
                            omega, fshift, flag, n)
    assert same_bits(back(dy, np.float64, n), hy)
    assert same_bits(back(di, np.float64, n), hi)


def _upload_factor(f, kind, n):
    nnz = int(f["bi"][n]) if kind == "icc" else int(f["bdiag"][0]) + 1
    out = {k: dev(f[k].astype(np.int32)) for k in ("bi", "bdiag", "r")}
    out["bj"] = dev(f["bj"][:max(nnz, 1)].astype(np.int32))
    out["ba"] = dev(f["ba"])
    return out


@pytest.mark.parametrize("name", sorted(MATS))
@pytest.mark.parametrize("kind,levels,ordering", [("ilu", 0, "natural"), ("ilu", 1, "natural"), ("ilu", 2, "rcm"),
                                                  ("icc", 0, "natural"), ("icc", 1, "natural"),
                                                  ("icc", 2, "natural"), ("icc", 3, "natural"),
                                                  ("icc", 0, "rcm"), ("icc", 1, "rcm"), ("icc", 2, "rcm")])
def test_factor_apply_bitwise(native, O, name, kind, levels, ordering):
    ip, ix, va = MATS[name]()
    if kind == "icc" and name in ("cd10", "rand500"):
        pytest.skip("IC needs a symmetric matrix")
    A = O.CSR(ip, ix, va)
    n = A.n
    pc = O.PCSetup(A, kind, levels, ordering)
    f = pc.factor
    b = rng_vec(n, 9)
    want = pc.apply(b)
    d = _upload_factor(f, kind, n)
    dx = dev(np.zeros(n))
    db = dev(b)  # keep the buffer alive across the call
    if kind == "ilu":
        native.mp_pc_ilu_apply(db.ptr, dx.ptr, 0, d["bi"].ptr, d["bj"].ptr, d["ba"].ptr, d["bdiag"].ptr,
                               d["r"].ptr, d["r"].ptr, n)
    else:
        native.mp_pc_icc_apply(db.ptr, dx.ptr, d["bi"].ptr, d["bj"].ptr, d["ba"].ptr, d["bdiag"].ptr,
                               d["r"].ptr, 0, n)
    assert same_bits(back(dx, np.float64, n), want)


def _rect(m, n, seed):
    """An m x n matrix with empty rows and columns no row touches."""
    rng = np.random.default_rng(seed)
    ip, ix = [0], []
    for i in range(m):
        cs = [] if i % 5 == 2 else sorted(set(rng.integers(0, max(n - 3, 1), int(rng.integers(1, 9))).tolist()))
        ix.extend(cs)
        ip.append(len(ix))
    return (np.array(ip, np.int64), np.array(ix, np.int64), rng.normal(size=len(ix)))


@pytest.mark.parametrize("shape", [(64, 64), (300, 170), (170, 300), (1, 5), (700, 700)])
def test_mat_surface_products_bitwise(O, shape):
    """Mat.mult / multTranspose / multTransposeAdd through the PETSc surface
    against the oracle (MatMult_SeqAIJ, MatMultTranspose_SeqAIJ), bit for bit,
    again after the values change (the sliced copies of A and A^T rebuilt)."""
    from mojopetsc4py import PETSc
    m, n = shape
    ip, ix, va = _rect(m, n, m * 7 + n)
    A = PETSc.Mat().createSeqAIJ((m, n), nnz=9)
    A.setValuesCSR(ip, ix, va)
    A.assemble()
    for rnd in range(2):
        xm, xn = rng_vec(m, 5 + rnd), rng_vec(n, 7 + rnd)
        ym = A.createVecLeft()
        A.mult(PETSc.Vec().createWithArray(xn), ym)
        want = np.zeros(m)
        O.lib().or_mat_mult(xn.ctypes.data, want.ctypes.data, ip.ctypes.data, ix.ctypes.data, va.ctypes.data, m)
        assert same_bits(ym.getArray(), want)
        yn = A.createVecRight()
        A.multTranspose(PETSc.Vec().createWithArray(xm), yn)
        want = np.zeros(n)
        O.lib().or_mat_mult_transpose(xm.ctypes.data, want.ctypes.data, ip.ctypes.data, ix.ctypes.data,
                                      va.ctypes.data, m, n)
        assert same_bits(yn.getArray(), want)
        v = rng_vec(n, 11)
        zn = A.multTransposeAdd(PETSc.Vec().createWithArray(xm), PETSc.Vec().createWithArray(v))
        assert same_bits(zn.getArray(), want + v)
        va = va * 1.5 - 0.25
        A.setValuesCSR(ip, ix, va)
        A.assemble()


ILU_DEV_MATS = dict(MATS)
ILU_DEV_MATS["cd24"] = lambda: W.convection_diffusion_3d(24)[0]
ILU_DEV_MATS["ragged300"] = lambda: _random_sparse(300, 14, 11)


@pytest.mark.parametrize("name", sorted(ILU_DEV_MATS))
@pytest.mark.parametrize("levels,ordering", [(0, "natural"), (1, "natural"), (2, "natural"), (0, "rcm"),
                                             (2, "rcm")])
def test_ilu_numeric_device_bitwise(native, O, name, levels, ordering):
    """MatILUFactorNumeric_SeqAIJ on the device (mpk_ilu_levels, then
    mpk_ilu_numeric level by level) gives the oracle's factor values bit for
    bit, over levels of fill and orderings."""
    ip, ix, va = ILU_DEV_MATS[name]()
    A = O.CSR(ip, ix, va)
    n = A.n
    f = O.build_factor(A, "ilu", levels, ordering)
    r, ic = O.permutations(A, ordering)
    nnz = int(f["bdiag"][0]) + 1
    lrow = np.zeros(n, np.int32)
    lptr = np.zeros(n + 1, np.int64)
    nlev = native.mpk_ilu_levels(f["bi"].ctypes.data, f["bj"].ctypes.data, f["bdiag"].ctypes.data, n,
                                  lrow.ctypes.data, lptr.ctypes.data)
    if nlev == -2:
        # a level-of-fill layout with rows lacking their own diagonal slot:
        # only the serial host pass reproduces it (the PC takes that path)
        bd = f["bdiag"]
        assert any(bd[i] <= bd[i + 1] or f["bj"][bd[i]] != i for i in range(n))
        return
    D = DevMat(ip, ix, va)
    d = {k: dev(np.asarray(v).astype(np.int32)) for k, v in
         (("r", r), ("ic", ic), ("bi", f["bi"]), ("bdiag", f["bdiag"]), ("bj", f["bj"][:nnz]), ("lrow", lrow))}
    dba = dev(np.full(nnz, np.nan))
    rc = native.mpk_ilu_numeric(n, D.rp.ptr, D.ci.ptr, D.v.ptr, d["r"].ptr, d["ic"].ptr, d["bi"].ptr,
                                d["bj"].ptr, d["bdiag"].ptr, d["lrow"].ptr, lptr.ctypes.data, nlev, dba.ptr)
    assert rc == 0
    assert same_bits(back(dba, np.float64, nnz), f["ba"][:nnz])


@pytest.mark.parametrize("mat,levels,ordering", [("cd16", 0, "natural"), ("cd16", 1, "rcm"),
                                                 ("ragged300", 1, "natural")])
def test_pc_ilu_set_up_on_device(O, mat, levels, ordering):
    """PC.setUp's ILU factor through the PETSc surface (computed on the
    device; ragged300 at level 1 has rows without their own diagonal slot and
    is factored by the serial host pass) equals the oracle's, and so does its
    apply."""
    from mojopetsc4py import PETSc
    ip, ix, va = W.convection_diffusion_3d(16)[0] if mat == "cd16" else ILU_DEV_MATS[mat]()
    n = len(ip) - 1
    A = PETSc.Mat().createSeqAIJ((n, n), nnz=7)
    A.setValuesCSR(ip, ix, va)
    A.assemble()
    pc = PETSc.PC().create()
    pc.setType("ilu")
    pc.setFactorLevels(levels)
    pc.setFactorOrdering(ordering)
    pc.setOperators(A)
    pc.setUp()
    want = O.build_factor(O.CSR(ip, ix, va), "ilu", levels, ordering)
    got = pc._host_factor()
    nnz = int(want["bdiag"][0]) + 1
    assert same_bits(got["ba"][:nnz], want["ba"][:nnz])
    b = rng_vec(n, 4)
    y = PETSc.Vec().createSeq(n)
    pc.apply(PETSc.Vec().createWithArray(b), y)
    assert same_bits(y.getArray(), O.PCSetup(O.CSR(ip, ix, va), "ilu", levels, ordering).apply(b))


ICC_DEV_MATS = {
    "lap64": lambda: W.laplacian_2d(64),
    "diff12": lambda: W.diffusion_3d(12)[0],
    "tri33": lambda: W.tridiagonal(33),
    "diff20": lambda: W.diffusion_3d(20)[0],
}


@pytest.mark.parametrize("name", sorted(ICC_DEV_MATS))
@pytest.mark.parametrize("levels,ordering", [(0, "natural"), (1, "natural"), (2, "natural"), (0, "rcm"),
                                             (2, "rcm")])
def test_icc_numeric_device_bitwise(native, O, name, levels, ordering):
    """MatICCFactorNumeric_SeqAIJ on the device (mpk_icc_visits' replay of
    the row chains, then mpk_icc_numeric level by level) gives the oracle's
    factor values bit for bit, including the layouts level-of-fill and the
    reordering give (columns below the diagonal, the unpermuted index in the
    diagonal slot)."""
    ip, ix, va = ICC_DEV_MATS[name]()
    A = O.CSR(ip, ix, va)
    n = A.n
    f = O.build_factor(A, "icc", levels, ordering)
    bi, bj = f["bi"], f["bj"]
    nnz = int(bi[n])
    vptr, vrow, vpos = np.zeros(n + 1, np.int32), np.zeros(nnz, np.int32), np.zeros(nnz, np.int32)
    lrow, lptr = np.zeros(n, np.int32), np.zeros(n + 1, np.int64)
    c = lambda x: x.ctypes.data
    nlev = native.mpk_icc_visits(c(bi), c(bj), c(A.irow), c(A.icol), c(f["r"]), c(f["rinv"]), n, c(vptr), c(vrow),
                                 c(vpos), c(lrow), c(lptr))
    assert nlev > 0
    D = DevMat(ip, ix, va)
    d = {k: dev(np.asarray(v).astype(np.int32)) for k, v in
         (("r", f["r"]), ("riip", f["rinv"]), ("bi", bi), ("bj", bj[:nnz]), ("vptr", vptr), ("vrow", vrow),
          ("vpos", vpos), ("lrow", lrow))}
    du0 = dev(np.zeros(nnz))
    dua = dev(np.full(nnz, np.nan))
    rc = native.mpk_icc_numeric(n, D.rp.ptr, D.ci.ptr, D.v.ptr, d["r"].ptr, d["riip"].ptr, d["bi"].ptr,
                                d["bj"].ptr, d["vptr"].ptr, d["vrow"].ptr, d["vpos"].ptr, d["lrow"].ptr, c(lptr),
                                nlev, du0.ptr, dua.ptr)
    assert rc == 0
    assert same_bits(back(dua, np.float64, nnz), f["ba"][:nnz])


@pytest.mark.parametrize("levels,ordering", [(0, "natural"), (2, "rcm")])
def test_pc_icc_set_up_on_device(O, levels, ordering):
    """PC.setUp's IC factor through the PETSc surface (computed on the
    device) equals the oracle's, and so does its apply."""
    from mojopetsc4py import PETSc
    ip, ix, va = W.diffusion_3d(16)[0]
    n = len(ip) - 1
    A = PETSc.Mat().createSeqAIJ((n, n), nnz=7)
    A.setValuesCSR(ip, ix, va)
    A.assemble()
    pc = PETSc.PC().create()
    pc.setType("icc")
    pc.setFactorLevels(levels)
    pc.setFactorOrdering(ordering)
    pc.setOperators(A)
    pc.setUp()
    want = O.build_factor(O.CSR(ip, ix, va), "icc", levels, ordering)
    got = pc._host_factor()
    nnz = int(want["bi"][n])
    assert same_bits(got["ba"][:nnz], want["ba"][:nnz])
    b = rng_vec(n, 4)
    y = PETSc.Vec().createSeq(n)
    pc.apply(PETSc.Vec().createWithArray(b), y)
    assert same_bits(y.getArray(), O.PCSetup(O.CSR(ip, ix, va), "icc", levels, ordering).apply(b))


def _edge_mats():
    """Small matrices for the branches of the numeric passes: one row; a
    stored zero below the diagonal (ILU skips a zero multiplier,
    ported.mojo:1003); a pivot that eliminates to exactly zero (1/0 = inf);
    and a 2 x 2 with a negative pivot."""
    one = (np.array([0, 1]), np.array([0]), np.array([4.0]))
    zmul = (np.array([0, 2, 5, 7]), np.array([0, 1, 0, 1, 2, 1, 2]),
            np.array([2.0, 1.0, 0.0, 3.0, 1.0, 1.0, 5.0]))
    # row 1: a10 = 2, a11 = 1 -> u11 = 1 - (2/2)*1 = 0 exactly
    zpiv = (np.array([0, 2, 4]), np.array([0, 1, 0, 1]), np.array([2.0, 1.0, 2.0, 1.0]))
    neg = (np.array([0, 2, 4]), np.array([0, 1, 0, 1]), np.array([-3.0, 1.0, 1.0, 2.0]))
    return {"one": one, "zmul": zmul, "zpiv": zpiv, "neg": neg}


@pytest.mark.parametrize("name", sorted(_edge_mats()))
@pytest.mark.parametrize("kind", ["ilu", "icc"])
def test_device_numeric_edge_cases(native, O, name, kind):
    """The device numeric passes on the edge cases give the oracle's factor
    bit for bit (inf where a pivot vanishes), through PC.setUp."""
    from mojopetsc4py import PETSc
    ip, ix, va = _edge_mats()[name]
    n = len(ip) - 1
    A = PETSc.Mat().createSeqAIJ((n, n), nnz=3)
    A.setValuesCSR(ip, ix, va)
    A.assemble()
    pc = PETSc.PC().create()
    pc.setType(kind)
    pc.setOperators(A)
    pc.setUp()
    want = O.build_factor(O.CSR(ip, ix, va), kind, 0, "natural")
    nnz = int(want["bi"][n]) if kind == "icc" else int(want["bdiag"][0]) + 1
    got = pc._host_factor()
    assert same_bits(got["ba"][:nnz], want["ba"][:nnz])

