"""The host C/C++ of the set-up path under AddressSanitizer and
UndefinedBehaviorSanitizer: mpk_factor.cpp (the product's symbolic and
numeric factorisations, level sets, IC visit chains and sweep planner) and
the oracle, built by tests/host_sanitize.mk with gcc, run through the host
set-up and planner tests in a child process (the sanitizer runtime must load
first, so the child starts with it preloaded).  Any out-of-bounds access,
use after free or undefined behaviour aborts the child with a report."""

import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _runtime(name):
    out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


@pytest.mark.timeout(900)
def test_host_set_up_under_sanitizers():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("no host compiler")
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("gcc has no AddressSanitizer runtime")
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "host_sanitize.mk")], check=True)
    out = os.path.join(HERE, "_san_build")
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["MPK_SAN_HOST_LIB"] = os.path.join(out, "libmpk_host_san.so")
    env["MPK_SAN_ORACLE_LIB"] = os.path.join(out, "liboracle_san.so")
    env["PYTHONPATH"] = os.pathsep.join([HERE, os.path.join(ROOT, "mojo-petsc4py_amd"), os.path.join(ROOT, "oracle"),
                                         env.get("PYTHONPATH", "")])
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "_san_plugin", "-p", "no:cacheprovider",
           "-m", "not gpu", os.path.join(HERE, "test_host_setup.py"), os.path.join(HERE, "test_sweep_plan.py")]
    proc = subprocess.run(cmd, capture_output=True, text=True, env=env, cwd=ROOT, timeout=850)
    text = proc.stdout + proc.stderr
    assert "AddressSanitizer" not in text and "runtime error:" not in text, text[-6000:]
    assert proc.returncode == 0, text[-6000:]
    assert " passed" in proc.stdout
