"""ASan + UBSan builds of the CPU code (tests/sanitize/Makefile), run on the CPU.

host_runtime_san: the library's host code (tree builder, planner, argument
checks, leaf host plans, error and cleanup paths; device calls fail without a
GPU) instrumented, driven by tests/sanitize/host_runtime_driver.cpp.
oracle_san: the CPU oracle, driven by tests/sanitize/oracle_driver.c.
A sanitizer report aborts the driver, so a clean exit with its OK line means
no invalid access, leak or undefined behaviour was seen on those paths.
"""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")

pytestmark = pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc") or shutil.which("make") is None,
                                reason="needs ROCm's hipcc and clang for the instrumented build")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-C", SAN, "-j8"], capture_output=True, text=True, timeout=1200)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return os.path.join(SAN, "_build")


def _run(path):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=1:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["HIP_VISIBLE_DEVICES"] = ""          # no device: the host paths and their error exits run
    r = subprocess.run([path], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-5000:]
    assert "SANITIZE_OK" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    return r.stdout


def test_host_runtime_under_asan_ubsan(built):
    print(_run(os.path.join(built, "host_runtime_san")))


def test_oracle_under_asan_ubsan(built):
    print(_run(os.path.join(built, "oracle_san")))
