"""The C++ CK mirror (sphinxsys_amd/host/sphx_ck.hpp) against the Python mirror.

Both drive the same library through the same C ABI calls on the same inputs,
so their results must be identical bit for bit: the dam-break geometry of the
C++ case header against the Python case generator, and the fluid state after
a run of advection steps (single-call driver and one CK object per stage)."""
import os
import subprocess

import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cpp") / "ck_mirror_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-Wall", "-o", exe,
                           os.path.join(ROOT, "tests", "cpp", "ck_mirror_check.cpp"), "-ldl"])
    return exe


def run_harness(exe, lib_path, dim, precision, steps, mode, prefix):
    out = subprocess.run([exe, lib_path, str(dim), str(precision), str(steps), mode, prefix],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    t, p, e, n = out.stdout.split()
    return float(t), float(p), float(e), int(n)


def python_run(lib, geom, precision, steps):
    sc = cases.DamBreak(lib, geom, precision)
    sc.initialize()
    for _ in range(steps):
        sc.advection_step()
    p, e = sc.observe()
    state = np.concatenate([sc.fluid.read("Position").ravel(), sc.fluid.read("Velocity").ravel(),
                            sc.fluid.read("Compression")])
    res = (sc.time, float(p[0]), e, sc.acoustic_steps)
    sc.close()
    return res, state


def check(exe, lib, lib_path, dim, precision, steps, tmp_path):
    geom = cases.dambreak_2d() if dim == 2 else cases.dambreak_3d(dp=0.04)
    ref, state = python_run(lib, geom, precision, steps)
    for mode in ("driver", "dynamics"):
        prefix = str(tmp_path / f"{mode}")
        got = run_harness(exe, lib_path, dim, precision, steps, mode, prefix)
        geo = np.fromfile(prefix + ".geometry.bin")
        nf, nw = geom.fluid_pos.size, geom.wall_pos.size
        assert np.array_equal(geo[:nf], geom.fluid_pos.ravel())
        assert np.array_equal(geo[nf:nf + nw], geom.wall_pos.ravel())
        assert np.array_equal(geo[nf + nw:], geom.wall_normal.ravel())
        assert got[0] == ref[0] and got[3] == ref[3], (mode, got, ref)
        assert got[1] == ref[1] and got[2] == ref[2], (mode, got, ref)
        assert np.array_equal(np.fromfile(prefix + ".state.bin"), state), mode


@pytest.mark.parametrize("dim", [2, 3])
def test_cpp_mirror_matches_python_mirror_on_oracle(harness, oracle, dim, tmp_path):
    check(harness, oracle, oracle.path, dim, capi.FP64, 30 if dim == 2 else 5, tmp_path)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_cpp_mirror_matches_python_mirror_on_gpu(harness, hip, precision, tmp_path):
    check(harness, hip, hip.path, 2, precision, 30, tmp_path)


@pytest.mark.gpu
def test_case_program_runs_on_gpu():
    exe = os.path.join(ROOT, "sphinxsys_amd", "lib", "dambreak")
    out = subprocess.run([exe, "--dim", "2", "--steps", "50", "--precision", "mixed"], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert "hip-gfx950" in out.stdout
    assert "particle-steps/s" in out.stdout.splitlines()[-1]
