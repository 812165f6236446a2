"""Parity of the HIP linear kernel correction and free-surface indication with
the oracle (through the C ABI, same inputs).  Gates: FP64 correction matrices,
forces and compression rates within 1e-10 relative L2, indicators bit-exact;
MIXED against the FP64 oracle with the tolerances written in each test."""
import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import cases, ck
from tests.test_gpu_parity import FP64_GATE, perturbed_pair, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("alpha", [0.0, 0.5])
def test_correction_matrix_fp64(hip, oracle, dim, alpha):
    pr = perturbed_pair(hip, oracle, dim)
    pr.both(lambda s: ck.LinearCorrectionMatrixComplex(s["inner"], *s["walls"], alpha=alpha).exec())
    g, c = pr.read("LinearCorrectionMatrix")
    print(f"dim={dim} alpha={alpha} B rel_l2={rel_l2(g, c):.2e}", flush=True)
    assert rel_l2(g, c) < FP64_GATE
    assert np.abs(c - np.eye(dim)).max() > 0.1  # surface and wall rows are corrected
    pr.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("riemann", [capi.RIEMANN_ACOUSTIC, capi.RIEMANN_NONE])
def test_corrected_acoustic_step_fp64(hip, oracle, dim, riemann):
    pr = perturbed_pair(hip, oracle, dim)
    pr.both(lambda s: ck.LinearCorrectionMatrixComplex(s["inner"], *s["walls"], alpha=0.5).exec())
    dt = 2e-5
    for sub in range(2):
        pr.both(lambda s: ck.AcousticStep1stHalf(s["inner"], *s["walls"], riemann=riemann,
                                                 correction=capi.CORRECTION_LINEAR).exec(dt))
        for name in ("Pressure", "Force", "CompressionRate", "Velocity", "Displacement"):
            g, c = pr.read(name)
            assert rel_l2(g, c) < FP64_GATE, (sub, name, rel_l2(g, c))
        pr.both(lambda s: ck.AcousticStep2ndHalf(s["inner"], *s["walls"], riemann=riemann,
                                                 correction=capi.CORRECTION_LINEAR).exec(dt))
        for name in ("Force", "CompressionRate", "Compression", "Density"):
            g, c = pr.read(name)
            assert rel_l2(g, c) < FP64_GATE, (sub, name, rel_l2(g, c))
    pr.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_corrected_step_with_written_general_matrices(hip, oracle, dim):
    """A caller-written, non-symmetric LinearCorrectionMatrix is used as given."""
    pr = perturbed_pair(hip, oracle, dim, seed=4)
    n = pr.gpu["f"].n
    rng = np.random.default_rng(17)
    pr.write("LinearCorrectionMatrix", np.eye(dim) + 0.2 * rng.normal(size=(n, dim, dim)))
    dt = 2e-5
    pr.both(lambda s: ck.AcousticStep1stHalf(s["inner"], *s["walls"],
                                             correction=capi.CORRECTION_LINEAR).exec(dt))
    pr.both(lambda s: ck.AcousticStep2ndHalf(s["inner"], *s["walls"],
                                             correction=capi.CORRECTION_LINEAR).exec(dt))
    for name in ("Velocity", "Force", "CompressionRate", "Compression"):
        g, c = pr.read(name)
        assert rel_l2(g, c) < FP64_GATE, (name, rel_l2(g, c))
    pr.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_corrected_step_mixed_against_fp64_oracle(hip, oracle, dim):
    """FP32 pair sums, update in double: B within 1e-4, forces within 5e-4 relative L2."""
    pr = perturbed_pair(hip, oracle, dim, precision=capi.MIXED)
    pr.both(lambda s: ck.LinearCorrectionMatrixComplex(s["inner"], *s["walls"], alpha=0.5).exec())
    g, c = pr.read("LinearCorrectionMatrix")
    assert rel_l2(g, c) < 1e-4, rel_l2(g, c)
    dt = 2e-5
    pr.both(lambda s: ck.AcousticStep1stHalf(s["inner"], *s["walls"],
                                             correction=capi.CORRECTION_LINEAR).exec(dt))
    for name, tol in (("Force", 5e-4), ("CompressionRate", 2e-4), ("Velocity", 1e-5)):
        g, c = pr.read(name)
        assert rel_l2(g, c) < tol, (name, rel_l2(g, c))
    pr.both(lambda s: ck.AcousticStep2ndHalf(s["inner"], *s["walls"],
                                             correction=capi.CORRECTION_LINEAR).exec(dt))
    for name, tol in (("Force", 5e-4), ("CompressionRate", 5e-4), ("Compression", 1e-6)):
        g, c = pr.read(name)
        assert rel_l2(g, c) < tol, (name, rel_l2(g, c))
    pr.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_surface_indication(hip, oracle, dim, precision):
    """Two calls: the first from PreviousSurfaceIndicator = 1, the second from a
    random previous pattern (exercises the temporal rule).  FP64: indicators
    bit-exact, divergence within 1e-10; MIXED: divergence within 1e-5 and at most
    0.5 % of the indicators differ (decisions within float rounding of the threshold)."""
    pr = perturbed_pair(hip, oracle, dim, precision=precision)
    n = pr.gpu["f"].n
    fp64 = precision == capi.FP64
    for call in range(2):
        if call == 1:
            pr.write("PreviousSurfaceIndicator", np.random.default_rng(3).integers(0, 2, n).astype(np.int32))
        pr.both(lambda s: ck.FreeSurfaceIndicationComplexSpatialTemporalCK(s["inner"], *s["walls"]).exec())
        g, c = pr.read("PositionDivergence")
        assert rel_l2(g, c) < (FP64_GATE if fp64 else 1e-5), (call, rel_l2(g, c))
        gi, ci = pr.read("Indicator")
        assert 0 < ci.sum() < n
        if fp64:
            assert np.array_equal(gi, ci), call
        else:
            assert np.count_nonzero(gi != ci) <= 0.005 * n, (call, np.count_nonzero(gi != ci))
        gp, cp = pr.read("PreviousSurfaceIndicator")
        assert np.array_equal(gp, gi)
    pr.close()


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_dambreak_2d_corrected_short_run(hip, oracle, precision):
    """25 advection steps of the 2D dam break with LinearCorrectionMatrix (alpha 0.5),
    the corrected half steps and the surface indication, a sort after the 20th,
    through the single-call driver.  FP64: same sub-step counts, state within 1e-9."""
    geom = cases.dambreak_2d(dp=0.025)
    kw = dict(correction=capi.CORRECTION_LINEAR, surface_indication=True)
    gpu = cases.DamBreak(hip, geom, precision, **kw)
    cpu = cases.DamBreak(oracle, geom, capi.FP64, **kw)
    gpu.initialize()
    cpu.initialize()
    for k in range(25):
        sort = k == 19
        gpu.time, ng = gpu.advection.exec(gpu.time, sort_now=sort)
        cpu.time, nc = cpu.advection.exec(cpu.time, sort_now=sort)
        if precision == capi.FP64:
            assert ng == nc, k
    fp64 = precision == capi.FP64
    assert gpu.time == pytest.approx(cpu.time, rel=1e-12 if fp64 else 2e-2)
    assert np.array_equal(gpu.fluid.read("OriginalID"), cpu.fluid.read("OriginalID"))
    tol = {"Position": 1e-12 if fp64 else 1e-4, "Velocity": 1e-9 if fp64 else 3e-2,
           "Compression": 1e-9 if fp64 else 3e-2, "LinearCorrectionMatrix": 1e-9 if fp64 else 1e-3}
    for name, t in tol.items():
        g, c = gpu.fluid.read(name), cpu.fluid.read(name)
        if name == "Compression":
            g, c = g - 1.0, c - 1.0
        assert rel_l2(g, c) < t, (name, rel_l2(g, c))
    gp, cp = gpu.fluid.read("PreviousSurfaceIndicator"), cpu.fluid.read("PreviousSurfaceIndicator")
    if fp64:
        assert np.array_equal(gp, cp)
    else:
        assert np.count_nonzero(gp != cp) <= 0.01 * len(cp)
    gpu.close()
    cpu.close()
