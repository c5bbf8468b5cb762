"""AddressSanitizer + UndefinedBehaviorSanitizer runs of the host C/C++ code.

Two executables are built by tests/sanitize/Makefile:
- oracle_check: the C oracle (oracle/feec_oracle.c) driven through every
  entry point on small Kuhn meshes in d = 1..4 (topology, metrics, element
  matrices of all pairings and grades, assembly, sparse algebra, CG, MINRES,
  Chebyshev MINRES);
- host_check: libfeec's translation units with the sanitizers on their host
  pass, driving the host-only entry points (meshers, option defaults, status
  strings, element shapes, argument checks including enum values outside
  their declarations) and the multigrid set-up's dense Cholesky inverse and
  tridiagonal eigenvalue bisection. No GPU is touched.
Any sanitizer report aborts the executable (-fno-sanitize-recover=all), and a
leak fails it at exit.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
HIPCC = "/opt/rocm/bin/hipcc"


def _env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:verify_asan_link_order=0:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["OMP_NUM_THREADS"] = "2"
    return env


@pytest.fixture(scope="module")
def built():
    if shutil.which("make") is None or shutil.which("gcc") is None:
        pytest.skip("make and gcc are needed for the sanitizer build")
    targets = ["_build/oracle_check"]
    if os.path.exists(HIPCC):
        targets.append("_build/host_check")
    jobs = str(min(8, os.cpu_count() or 2))
    res = subprocess.run(["make", "-s", "-j", jobs, *targets], cwd=SAN, capture_output=True,
                         text=True, timeout=900)
    assert res.returncode == 0, res.stdout + res.stderr
    return targets


def _run(exe):
    res = subprocess.run([os.path.join(SAN, exe)], cwd=SAN, capture_output=True, text=True,
                         env=_env(), timeout=600)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out
    assert "ERROR: LeakSanitizer" not in out, out
    return out


def test_oracle_under_asan_ubsan(built):
    assert "oracle_check ok" in _run("_build/oracle_check")


def test_libfeec_host_code_under_asan_ubsan(built):
    if "_build/host_check" not in built:
        pytest.skip("hipcc is needed to build libfeec's translation units")
    assert "host_check ok" in _run("_build/host_check")
