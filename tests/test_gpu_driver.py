"""The single-call advection step against the same step written out with one
dynamics object per stage.

sphx_advection_step runs the acoustic sub-steps with their time step and loop
condition on the device, the 2nd half's initialize fused into pass 1 and the
1st half's force left unstored inside the loop (the 2nd half sets it anew).
The written-out step (cases.DamBreak.advection_step_by_dynamics, as
dambreak.cpp:183-233 lists its dynamics) computes each time step on the host
and launches every half on its own.  After the steps both give the same
sub-step counts and, within the stated tolerances, the same state, the force
included: FP64 within 1e-10 relative L2 (the two time-step paths may round
their last bit differently), FP32-mixed within 1e-5."""
import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import cases

pytestmark = pytest.mark.gpu

FIELDS = ("Position", "Velocity", "Force", "Density", "Pressure", "Compression", "CompressionRate")


def rel_l2(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("dim,precision", [(2, capi.FP64), (3, capi.FP64), (3, capi.MIXED)])
def test_single_call_step_matches_written_out_step(hip, dim, precision):
    geom = cases.dambreak_2d(dp=0.05) if dim == 2 else cases.dambreak_3d(dp=0.04)
    one = cases.DamBreak(hip, geom, precision)
    two = cases.DamBreak(hip, geom, precision)
    one.initialize()
    two.initialize()
    print(f"{dim}D dam break, {one.fluid.n} fluid particles, precision {precision}", flush=True)
    tol = 1e-10 if precision == capi.FP64 else 1e-5
    try:
        for k in range(30):
            n_one = one.advection_step()
            n_two = two.advection_step_by_dynamics()
            assert n_one == n_two, (k, n_one, n_two)
        assert one.time == pytest.approx(two.time, rel=1e-12)
        assert np.array_equal(one.fluid.read("OriginalID"), two.fluid.read("OriginalID"))
        for name in FIELDS:
            err = rel_l2(one.fluid.read(name), two.fluid.read(name))
            assert err <= tol, (name, err)
    finally:
        one.close()
        two.close()
