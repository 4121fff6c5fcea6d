"""The device IC(0) and ILU(0) analyses against the host passes they
replace, array for array: mpk_ic0_analyse against mp_icc_symbolic at level 0
with mp_icc_contiguous_cholesky, mpk_icc_transpose and mpk_icc_visits;
mpk_ilu0_analyse against mp_ilu_symbolic_ilu0 and mpk_ilu_levels; the
matrices they decline; and PC.setUp through them against the host passes
(MPK_IC0_DEVICE=0, MPK_ILU0_DEVICE=0) and the oracle, bit for bit."""

import numpy as np
import pytest

from _helpers import back, dev, rng_vec, same_bits
from mpk import workloads as W

pytestmark = pytest.mark.gpu


def _a(x):
    return x.ctypes.data


def _random_pattern(n, density, seed, symmetric, diag):
    rng = np.random.default_rng(seed)
    M = rng.random((n, n)) < density
    if symmetric:
        M |= M.T
    np.fill_diagonal(M, diag)
    ip = np.concatenate([[0], np.cumsum(M.sum(1))]).astype(np.int64)
    ix = np.nonzero(M)[1].astype(np.int64)
    va = rng.normal(size=len(ix))
    va[ix == np.repeat(np.arange(n), np.diff(ip))] += 3.0 * n * density + 4.0
    return ip, ix, va


MATS = {
    "one": lambda: (np.array([0, 1]), np.array([0]), np.array([4.0])),
    "tri40": lambda: W.tridiagonal(40),
    "lap30": lambda: W.laplacian_2d(30),
    "diff10": lambda: W.diffusion_3d(10)[0],
    "cd9": lambda: W.convection_diffusion_3d(9)[0],
    "sym200": lambda: _random_pattern(200, 0.05, 1, True, True),
    "nonsym150": lambda: _random_pattern(150, 0.06, 2, False, True),
    "nodiag120": lambda: _random_pattern(120, 0.08, 3, True, False),
    "nonsym_nodiag90": lambda: _random_pattern(90, 0.1, 4, False, False),
    "spd5k": lambda: W.random_spd(5000)[0],
}


def _host_analysis(L, ip, ix, n):
    """The host passes, as PC.setUp ran them before the device analysis."""
    cap = int(ip[-1]) + n + 16
    ui, uj, ud, lv = (np.zeros(k, np.int64) for k in (n + 1, cap, n, cap))
    sc = np.zeros(6 * (n + 1) + 2 * cap, np.int64)
    r = np.arange(n, dtype=np.int64)
    nnz = L.mp_icc_symbolic(_a(ui), _a(uj), _a(lv), _a(ud), _a(ip), _a(ix), _a(r), _a(r), 0, cap, _a(sc), n,
                            int(sc.size))
    assert nnz > 0
    bi = np.zeros(n + 1, np.int64)
    bd = np.zeros(n + 1, np.int64)
    bd[:n] = ud
    bj = np.zeros(cap, np.int64)
    L.mp_icc_contiguous_cholesky(_a(uj), _a(ui), _a(bd), _a(bi), _a(bd), _a(bj), n)
    vptr, vrow, vpos, lrow = (np.zeros(k, np.int32) for k in (n + 1, nnz, nnz, n))
    lptr = np.zeros(n + 1, np.int64)
    nlev = L.mpk_icc_visits(_a(bi), _a(bj), _a(ip), _a(ix), _a(r), _a(r), n, _a(vptr), _a(vrow), _a(vpos),
                            _a(lrow), _a(lptr))
    assert nlev > 0
    dbi, dbj, dbd = dev(bi.astype(np.int32)), dev(bj[:nnz].astype(np.int32)), dev(bd.astype(np.int32))
    m = L.mpk_icc_transpose_size(dbi.ptr, dbd.ptr, n)
    t = [dev(np.zeros(max(k, 1), np.int32)) for k in (n + 1, m, m, n)]
    nlate = L.mpk_icc_transpose(dbi.ptr, dbj.ptr, dbd.ptr, n, *(b.ptr for b in t))
    assert nlate == 0
    trp, trow, tpos, tmid = (back(b, np.int32, k) for b, k in zip(t, (n + 1, m, m, n)))
    return dict(nnz=nnz, bi=bi, bj=bj[:nnz], bdiag=bd, trp=trp, trow=trow, tpos=tpos, tmid=tmid, vptr=vptr,
                vrow=vrow[:vptr[n]], vpos=vpos[:vptr[n]], lrow=lrow, lptr=lptr[:nlev + 1], nlev=nlev)


def _device_analysis(L, ip, ix, n):
    rp, ci = dev(np.asarray(ip, np.int32)), dev(np.asarray(ix, np.int32))
    nnz = L.mpk_ic0_size(rp.ptr, ci.ptr, n)
    assert nnz > 0, nnz
    m = max(nnz - n, 1)
    sizes = dict(bi=n + 1, bj=nnz, bdiag=n + 1, trp=n + 1, trow=m, tpos=m, tmid=n, vrow=m, vpos=m, lrow=n)
    bufs = {k: dev(np.full(v, -7, np.int32)) for k, v in sizes.items()}
    lptr = np.zeros(n + 1, np.int64)
    nlev = L.mpk_ic0_analyse(rp.ptr, ci.ptr, n, *(bufs[k].ptr for k in sizes), _a(lptr))
    assert nlev > 0
    out = {k: back(bufs[k], np.int32, v) for k, v in sizes.items()}
    out.update(nnz=nnz, nlev=nlev, lptr=lptr[:nlev + 1])
    ms = nnz - n
    out["trow"], out["tpos"], out["vrow"], out["vpos"] = (out[k][:ms] for k in ("trow", "tpos", "vrow", "vpos"))
    return out


@pytest.mark.parametrize("name", sorted(MATS))
def test_device_analysis_matches_host(native, name):
    ip, ix, _ = MATS[name]()
    ip, ix = np.asarray(ip, np.int64), np.asarray(ix, np.int64)
    n = len(ip) - 1
    want = _host_analysis(native, ip, ix, n)
    got = _device_analysis(native, ip, ix, n)
    assert got["nnz"] == want["nnz"] and got["nlev"] == want["nlev"]
    for key in ("bi", "bj", "bdiag", "trp", "trow", "tpos", "tmid", "vrow", "vpos", "lrow", "lptr"):
        assert np.array_equal(got[key], want[key]), key
    assert np.array_equal(got["trp"], want["vptr"])  # the visits start where the columns do


def test_declined_matrices(native):
    """Unsorted or repeated columns, and a column of over 256 entries, are
    declined with -2 (PC.setUp then takes the host passes)."""
    L = native

    def size(ip, ix):
        rp, ci = dev(np.asarray(ip, np.int32)), dev(np.asarray(ix, np.int32))  # alive through the call
        return L.mpk_ic0_size(rp.ptr, ci.ptr, len(ip) - 1)

    assert size([0, 2, 3], [1, 0, 1]) == -2          # row 0 descending
    assert size([0, 3, 4], [0, 1, 1, 1]) == -2       # a repeated column
    assert size([0, 2, 3], [0, 5, 1]) == -2          # a column out of range
    n = 300                                           # the last column dense
    ip = np.arange(n + 1) * 2
    ip[n] = 2 * n - 1
    ix = np.stack([np.arange(n), np.full(n, n - 1)], 1).ravel()[:2 * n - 1]
    assert size(ip, ix) == -2
    m = 200                                           # 199 entries: accepted
    ip = np.arange(m + 1) * 2
    ip[m] = 2 * m - 1
    ix = np.stack([np.arange(m), np.full(m, m - 1)], 1).ravel()[:2 * m - 1]
    assert size(ip, ix) == 2 * (m - 1) + 1


def _pc(csr, monkeypatch, device, kind="icc"):
    from mojopetsc4py import PETSc
    switch = "MPK_IC0_DEVICE" if kind == "icc" else "MPK_ILU0_DEVICE"
    if device:
        monkeypatch.delenv(switch, raising=False)
    else:
        monkeypatch.setenv(switch, "0")
    ip, ix, va = csr
    n = len(ip) - 1
    A = PETSc.Mat().createSeqAIJ((n, n), nnz=7)
    A.setValuesCSR(ip, ix, va)
    A.assemble()
    pc = PETSc.PC().create()
    pc.setType(kind)
    pc.setOperators(A)
    pc.setUp()
    return pc


@pytest.mark.parametrize("name", ["tri40", "diff10", "sym200", "nonsym150", "nodiag120", "spd5k"])
def test_pc_setup_through_device_analysis(native, oracle, monkeypatch, name):
    """PC.setUp's IC(0) through the device analysis: the factor, the apply
    and the factor matrix bit for bit those of the host passes, and the
    factor the oracle's."""
    from mojopetsc4py import PETSc
    csr = MATS[name]()
    n = len(csr[0]) - 1
    on = _pc(csr, monkeypatch, True)
    assert on._factor["bi"] is None  # the device path ran: nothing on the host yet
    off = _pc(csr, monkeypatch, False)
    f_on, f_off = on._host_factor(), off._host_factor()
    nnz = int(f_off["bi"][n])
    for key in ("bi", "bdiag"):
        assert np.array_equal(f_on[key], f_off[key]), key
    assert np.array_equal(f_on["bj"][:nnz], f_off["bj"][:nnz])
    assert same_bits(f_on["ba"][:nnz], f_off["ba"][:nnz])
    want = oracle.build_factor(oracle.CSR(*csr), "icc", 0, "natural")
    # bit for bit, but for NaN payloads: the device's default NaN has another
    # sign bit than the host's (a pattern without diagonals divides by 0)
    got, ref = f_on["ba"][:nnz], want["ba"][:nnz]
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    keep = ~np.isnan(ref)
    assert same_bits(got[keep], ref[keep])
    b = rng_vec(n, 9)
    ys = []
    for pc in (on, off):
        y = PETSc.Vec().createSeq(n)
        pc.apply(PETSc.Vec().createWithArray(b), y)
        ys.append(y.getArray().copy())
    assert same_bits(ys[0], ys[1])
    if n <= 1000:  # the factor matrix goes through a dense product
        assert same_bits(on.getFactorMatrix().getValuesCSR()[2], off.getFactorMatrix().getValuesCSR()[2])


ILU_MATS = ["one", "tri40", "lap30", "diff10", "cd9", "sym200", "nonsym150", "spd5k"]


def _host_ilu0(L, ip, ix, n):
    adiag = np.array([ip[i] + int(np.searchsorted(ix[ip[i]:ip[i + 1]], i)) for i in range(n)], np.int64)
    bi, bd = np.zeros(n + 1, np.int64), np.zeros(n + 1, np.int64)
    bj = np.zeros(int(ip[-1]) + n + 16, np.int64)
    nnz = L.mp_ilu_symbolic_ilu0(_a(bi), _a(bj), _a(bd), _a(adiag), _a(ip), _a(ix), n)
    lrow, lptr = np.zeros(n, np.int32), np.zeros(n + 1, np.int64)
    nlev = L.mpk_ilu_levels(_a(bi), _a(bj), _a(bd), n, _a(lrow), _a(lptr))
    assert nlev > 0
    return dict(nnz=nnz, bi=bi, bj=bj[:nnz], bdiag=bd, lrow=lrow, lptr=lptr[:nlev + 1], nlev=nlev)


@pytest.mark.parametrize("name", ILU_MATS)
def test_device_ilu0_analysis_matches_host(native, name):
    L = native
    ip, ix, _ = MATS[name]()
    ip, ix = np.asarray(ip, np.int64), np.asarray(ix, np.int64)
    n = len(ip) - 1
    want = _host_ilu0(L, ip, ix, n)
    rp, ci = dev(ip.astype(np.int32)), dev(ix.astype(np.int32))
    nnz = L.mpk_ilu0_size(rp.ptr, ci.ptr, n)
    assert nnz == want["nnz"] == want["bdiag"][0] + 1
    sizes = dict(bi=n + 1, bj=nnz, bdiag=n + 1, lrow=n)
    bufs = {k: dev(np.full(v, -7, np.int32)) for k, v in sizes.items()}
    lptr = np.zeros(n + 1, np.int64)
    nlev = L.mpk_ilu0_analyse(rp.ptr, ci.ptr, n, *(bufs[k].ptr for k in sizes), _a(lptr))
    assert nlev == want["nlev"]
    assert np.array_equal(lptr[:nlev + 1], want["lptr"])
    for key, v in sizes.items():
        assert np.array_equal(back(bufs[key], np.int32, v), want[key]), key


def test_ilu0_declined_matrices(native):
    L = native

    def size(ip, ix):
        rp, ci = dev(np.asarray(ip, np.int32)), dev(np.asarray(ix, np.int32))  # alive through the call
        return L.mpk_ilu0_size(rp.ptr, ci.ptr, len(ip) - 1)

    assert size([0, 2, 3], [1, 0, 1]) == -2          # row 0 descending
    assert size([0, 2, 3], [0, 1, 0]) == -2          # row 1 without its diagonal
    assert size([0, 2, 3], [0, 4, 1]) == -2          # a column out of range
    assert size([0, 2, 4], [0, 1, 0, 1]) == 4


@pytest.mark.parametrize("name", ["tri40", "diff10", "cd9", "nonsym150", "spd5k"])
def test_pc_ilu_setup_through_device_analysis(native, oracle, monkeypatch, name):
    """PC.setUp's ILU(0) through the device analysis: the factor and the
    apply bit for bit those of the host passes, the factor the oracle's."""
    from mojopetsc4py import PETSc
    csr = MATS[name]()
    n = len(csr[0]) - 1
    on = _pc(csr, monkeypatch, True, "ilu")
    assert on._factor["bi"] is None
    off = _pc(csr, monkeypatch, False, "ilu")
    f_on, f_off = on._host_factor(), off._host_factor()
    nnz = int(f_off["bdiag"][0]) + 1
    for key in ("bi", "bdiag"):
        assert np.array_equal(f_on[key], f_off[key]), key
    assert np.array_equal(f_on["bj"][:nnz], f_off["bj"][:nnz])
    assert same_bits(f_on["ba"][:nnz], f_off["ba"][:nnz])
    want = oracle.build_factor(oracle.CSR(*csr), "ilu", 0, "natural")
    assert same_bits(f_on["ba"][:nnz], want["ba"][:nnz])
    b = rng_vec(n, 10)
    ys = []
    for pc in (on, off):
        y = PETSc.Vec().createSeq(n)
        pc.apply(PETSc.Vec().createWithArray(b), y)
        ys.append(y.getArray().copy())
    assert same_bits(ys[0], ys[1])
    if n <= 1000:
        assert same_bits(on.getFactorMatrix().getValuesCSR()[2], off.getFactorMatrix().getValuesCSR()[2])
