"""State I/O across implementations.

* A restart snapshot written from the device state (2D dam break, linear
  correction and surface indication on, FP64, after 20 advection steps, the
  reference's evolving variables plus "Force") is resumed on the oracle; 10
  advection steps later the oracle matches its own uninterrupted 30-step run
  (Position within 1e-10, Velocity within 1e-8 relative L2: the snapshot
  prints 15 digits, the two sides sum in different orders).
* VTP files written from the device state and from the oracle state after
  the same stages on the same FP64 input (summation, regularization, surface
  indication on the initial lattice of the reduced 3D dam break) are equal
  byte for byte.
"""
import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import cases, state_io

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make(lib, geom):
    return cases.DamBreak(lib, geom, capi.FP64, correction=capi.CORRECTION_LINEAR, surface_indication=True)


def test_device_restart_resumes_on_the_oracle(hip, oracle, tmp_path):
    geom = cases.dambreak_2d(dp=0.025)
    gpu = make(hip, geom)
    gpu.initialize()
    for _ in range(20):
        gpu.advection_step()
    evolving = state_io.FLUID_EVOLVING + [("Force", "vec")]
    state_io.RestartIOCK(str(tmp_path), [("WaterBody", gpu.fluid)], {"WaterBody": evolving}).write_to_file(
        20, gpu.time)
    gpu.close()

    ref = make(oracle, geom)
    ref.initialize()
    for _ in range(30):
        ref.advection_step()

    res = make(oracle, geom)
    io = state_io.RestartIOCK(str(tmp_path), [("WaterBody", res.fluid)], {"WaterBody": evolving})
    res.time = io.read_from_file(20)
    res.iterations = 20
    # configuration from the restored positions (dambreak_sycl.cpp:137-142 on restart)
    res.fluid_cell_linked_list.exec()
    res.wall_cell_linked_list.exec()
    res.complex_relation.exec()
    for _ in range(10):
        res.advection_step()
    ep = rel_l2(res.fluid.read("Position"), ref.fluid.read("Position"))
    ev = rel_l2(res.fluid.read("Velocity"), ref.fluid.read("Velocity"))
    print(f"device restart resumed on the oracle: Position {ep:.2e}, Velocity {ev:.2e}", flush=True)
    assert res.time == pytest.approx(ref.time, rel=1e-12)
    assert ep < 1e-10 and ev < 1e-8
    assert np.array_equal(res.fluid.read("OriginalID"), ref.fluid.read("OriginalID"))
    res.close()
    ref.close()


def test_vtp_bytes_equal_between_device_and_oracle(hip, oracle, tmp_path):
    geom = cases.dambreak_3d(dp=0.05)
    paths = []
    for name, lib in (("gpu", hip), ("cpu", oracle)):
        sc = make(lib, geom)
        sc.initialize()
        sc.summation.exec()
        sc.regularization.exec()
        sc.boundary_indicator.exec()
        (tmp_path / name).mkdir()
        rec = state_io.BodyStatesRecordingToVtp(str(tmp_path / name), [("WaterBody", sc.fluid)])
        rec.add_to_write("WaterBody", "Density", "real")
        rec.add_to_write("WaterBody", "Indicator", "int")
        rec.add_to_write("WaterBody", "PositionDivergence", "real")
        rec.add_to_write("WaterBody", "CompressionSummation", "real")
        paths.append(rec.write_to_file(0.0, iteration=0)[0])
        sc.close()
    a, b = (open(p, "rb").read() for p in paths)
    assert len(a) > 10000
    assert a == b
