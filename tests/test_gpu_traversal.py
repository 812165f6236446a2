"""The second acoustic half step with the inner neighbours found by a cell
traversal (sphx_acoustic_step_2nd_half_traversal) against the same step on the
stored lists: the traversal takes the neighbours in list order by the list
build's criterion, so every output is bit-identical.

The scene is the 3D dam break at dp = 0.03 in FP32-mixed after 60 advection
steps (particles out of their sorted order, the free surface moving), once
with the default lists and once with lists from a scene advanced the same way;
both scenes are deterministic, so their states are bit-identical before the
compared call."""
import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import cases, ck

pytestmark = pytest.mark.gpu

OUTPUTS = ("Force", "CompressionRate", "Compression", "Density", "Displacement")


def advanced_scene(lib, geom, steps):
    sc = cases.DamBreak(lib, geom, capi.MIXED)
    sc.initialize()
    for _ in range(steps):
        sc.advection_step()
    sc.ctx.synchronize()
    return sc


@pytest.mark.parametrize("riemann", [capi.RIEMANN_ACOUSTIC, capi.RIEMANN_NONE])
def test_traversal_equals_stored_lists(hip, riemann):
    geom = cases.dambreak_3d(dp=0.03)
    a = advanced_scene(hip, geom, 60)
    b = advanced_scene(hip, geom, 60)
    for k in OUTPUTS + ("Velocity", "Position"):
        assert np.array_equal(a.fluid.read(k), b.fluid.read(k)), k
    dt = 2.0e-4
    ck.AcousticStep2ndHalf(a.inner, a.contact, riemann=riemann).exec(dt)
    ck.AcousticStep2ndHalfTraversal(b.inner, b.contact, riemann=riemann).exec(dt)
    a.ctx.synchronize()
    b.ctx.synchronize()
    for k in OUTPUTS:
        fa, fb = a.fluid.read(k), b.fluid.read(k)
        assert np.all(np.isfinite(fa)), k
        assert np.array_equal(fa, fb), (k, float(np.abs(fa - fb).max()))
    # the step did work: the velocity divergence of a falling column is not zero
    # (without a Riemann solver the second half adds no force)
    assert np.abs(a.fluid.read("CompressionRate")).max() > 0.0
    if riemann != capi.RIEMANN_NONE:
        assert np.abs(a.fluid.read("Force")).max() > 0.0
    a.close()
    b.close()


def test_traversal_rejects_fp64(hip):
    geom = cases.dambreak_3d(dp=0.05)
    sc = cases.DamBreak(hip, geom, capi.FP64)
    sc.initialize()
    with pytest.raises(capi.SphxError, match="FP32-mixed"):
        ck.AcousticStep2ndHalfTraversal(sc.inner, sc.contact).exec(1e-4)
    sc.close()
