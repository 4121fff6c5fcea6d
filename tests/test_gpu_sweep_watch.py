"""Bounded sweep waits.  Every wait of a device sweep for a row another
workgroup publishes gives up after a poll limit, or once the host raises the
abort word, and its kernel then runs to its end.  With a limit of one poll
(or the abort word raised) an apply either finds every row it waits for
already published -- and gives the oracle's bits with no error -- or some
wait gives up and the library reports status 10 at its next
synchronisation; either way the kernels end and the next apply, back at the
default limit, gives the oracle's bits again.  Sweeps of block plans, ring
plans and wave blocks (multi-workgroup, with published reads) are covered."""

import numpy as np
import pytest

from _helpers import rng_vec, same_bits
from mpk import workloads as W

pytestmark = pytest.mark.gpu

MATS = {
    "diff24": lambda: W.diffusion_3d(24)[0],
    "lap100": lambda: W.laplacian_2d(100),
    "spd20k": lambda: W.random_spd(20000, band=3000)[0],
}
CASES = [("diff24", "icc", "auto", "default"), ("diff24", "ilu", "blocks", "0"), ("lap100", "ilu", "ring", "0"),
         ("spd20k", "sor", "auto", "default"), ("spd20k", "sor", "blocks", "0")]


def _setup(name, kind):
    from mojopetsc4py import PETSc
    ip, ix, va = MATS[name]()
    n = len(ip) - 1
    A = PETSc.Mat().createSeqAIJ((n, n), nnz=7)
    A.setValuesCSR(ip, ix, va)
    A.assemble()
    pc = PETSc.PC().create()
    pc.setType(kind)
    if kind == "sor":
        pc.setSOROmega(1.2)
    pc.setOperators(A)
    pc.setUp()
    return (ip, ix, va), A, pc


def _apply(pc, b):
    from mojopetsc4py import PETSc
    from mpk._native import NativeError
    y = PETSc.Vec().createSeq(len(b))
    try:
        pc.apply(PETSc.Vec().createWithArray(b), y)
        return y.getArray().copy(), None
    except NativeError as e:
        return None, str(e)


@pytest.mark.parametrize("name,kind,plan,narrow", CASES, ids=["-".join(c) for c in CASES])
@pytest.mark.parametrize("how", ["limit1", "abort"])
def test_waits_give_up_and_kernels_end(native, oracle, monkeypatch, name, kind, plan, narrow, how):
    if plan != "auto":
        monkeypatch.setenv("MPK_SWEEP_PLAN", plan)
    if narrow == "0":
        monkeypatch.setenv("MPK_SWEEP_NARROW", "0")
    oracle.set_mode("dev")
    try:
        csr, A, pc = _setup(name, kind)
        n = A.getSize()[0]
        Ao = oracle.CSR(*csr)
        po = oracle.PCSetup(Ao, "sor", omega=1.2) if kind == "sor" else oracle.PCSetup(Ao, kind, 0, "natural")
        b = rng_vec(n, 7)
        want = po.apply(b)
        print(f"[watch] {name} {kind} plan={plan} narrow={narrow} {how}", flush=True)
        if how == "limit1":
            old = native.mpk_set_sweep_wait_limit(1)
        else:
            old = None
            assert native.mpk_sweep_abort(1) == 0
        try:
            gave = 0
            for _ in range(3):
                y, err = _apply(pc, b)
                native.mpk_synchronize()
                if err is None and native.mpk_status() != 0:
                    err = "status %d" % native.mpk_status()
                    native.mpk_clear_error()
                if err is None:
                    assert same_bits(y, want)
                else:
                    assert "gave up" in err, err
                    gave += 1
            print(f"[watch] {gave} of 3 applies gave up", flush=True)
        finally:
            if old is not None:
                native.mpk_set_sweep_wait_limit(old)
            native.mpk_sweep_abort(0)
            native.mpk_clear_error()
        y, err = _apply(pc, b)
        assert err is None, err
        assert same_bits(y, want)
        native.mpk_synchronize()
        assert native.mpk_status() == 0
    finally:
        oracle.set_mode("ref")
