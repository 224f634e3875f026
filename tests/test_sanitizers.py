"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU): the oracle's C
restatement (gcc build, loaded by the oracle tests with the ASan runtime preloaded) and the
Chebyshev store's host builders (spectral radius, series, preorder slot layout) in a harness
built by hipcc with the sanitizers on the host side only (tests/sanitize/)."""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
BUILD = os.path.join(SAN, "build")


def _make(target):
    r = subprocess.run(["make", "-C", SAN, os.path.join(BUILD, target)], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return os.path.join(BUILD, target)


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc"), reason="hipcc not installed")
def test_host_builders_under_sanitizers():
    exe = _make("layout_harness")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
                                UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1"))
    print(r.stdout[-500:], r.stderr[-3000:])
    assert r.returncode == 0 and "0 failures" in r.stdout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not installed")
def test_oracle_under_sanitizers():
    lib = _make("liboracle_san.so")
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    ubsan = subprocess.run(["gcc", "-print-file-name=libubsan.so"], capture_output=True,
                           text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=f"{asan}:{ubsan}", NEUWON_ORACLE_LIB=lib,
               ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider", os.path.join(HERE, "test_oracle.py")],
                       capture_output=True, text=True, timeout=1200, env=env,
                       cwd=os.path.dirname(HERE))
    print(r.stdout[-1500:], r.stderr[-3000:])
    assert r.returncode == 0
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
