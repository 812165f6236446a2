"""The oracle (the parity checker) under AddressSanitizer + UndefinedBehaviorSanitizer
and under ThreadSanitizer: the C++ CK-mirror harness and the oracle built with
the sanitizers (oracle/Makefile targets asan, tsan; host code only) run the 2D
and 3D dam break on the oracle's thread pool, and no report may appear.  The
sanitized results equal the plain build's (same arithmetic)."""
import os
import subprocess

import pytest

from tests._util import ROOT

BUILD = os.path.join(ROOT, "oracle", "build")


@pytest.fixture(scope="module")
def sanitized():
    subprocess.run(["make", "-s", "-j4", "-C", os.path.join(ROOT, "oracle"), "asan", "tsan"], check=True,
                   timeout=900)
    return BUILD


def run(exe, lib, dim, steps, mode, tmp_path, env):
    out = subprocess.run([os.path.join(BUILD, exe), os.path.join(BUILD, lib), str(dim), "0", str(steps), mode,
                          str(tmp_path / f"{exe}_{dim}_{mode}")],
                         capture_output=True, text=True, timeout=600,
                         env={**os.environ, "SPHX_ORACLE_THREADS": "4", **env})
    assert out.returncode == 0, out.stderr[-4000:]
    assert "Sanitizer" not in out.stderr and "runtime error" not in out.stderr, out.stderr[-4000:]
    return out.stdout.split()


@pytest.mark.parametrize("dim,steps", [(2, 10), (3, 2)])
def test_oracle_under_asan_ubsan(sanitized, dim, steps, tmp_path):
    env = {"ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1", "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}
    for mode in ("driver", "dynamics"):
        run("ck_mirror_check_asan", "libsphx_oracle_asan.so", dim, steps, mode, tmp_path, env)


@pytest.mark.parametrize("dim,steps", [(2, 4), (3, 1)])
def test_oracle_under_tsan(sanitized, dim, steps, tmp_path):
    run("ck_mirror_check_tsan", "libsphx_oracle_tsan.so", dim, steps, "driver", tmp_path,
        {"TSAN_OPTIONS": "halt_on_error=1"})
