"""The 3D dam-break workload (tank 3.2 x 1.0 x 1.5, column 1.2 x 1.0 x 0.6, three
wall layers, rho0 1000, g 9.81) on the HIP library against the oracle.

* Reduced resolution (dp = 0.02, 90k fluid particles): 25 advection steps of
  the single-call driver with a ParticleSortCK after the 20th, the loop of
  tests/tests_sycl/3d_examples/test_3d_dambreak_sycl/dambreak.cpp:183-233.
  FP64: the same sub-step counts, OriginalID equal, Position within 1e-12 and
  Velocity / Compression within 1e-9 relative L2; probe pressures inside the
  water within 1e-9.  FP32-mixed: the float tolerances written in the test.
* Full resolution (dp = 0.00448, 8,008,376 fluid + 2,884,710 wall particles):
  the initial inner and wall-contact lists bit-exact, the lattice's interior
  count of 80 neighbours in closed form, both time steps in closed form, then
  one advection step in FP64 (state within 1e-9, rebuilt lists bit-exact) and
  in FP32-mixed (float tolerances, lists bit-exact).
"""
import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import cases, ck

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def in_water_probes(geom):
    """Observer points inside the initial water column and at the foot of the
    end wall (zero until the surge arrives)."""
    LL, LW, LH = geom.meta["LL"], geom.meta["LW"], geom.meta["LH"]
    return np.array([[0.1 * LL, 0.5 * LW, 0.1 * LH], [0.5 * LL, 0.3 * LW, 0.5 * LH],
                     [0.9 * LL, 0.8 * LW, 0.2 * LH], [geom.meta["DL"], 0.5 * LW, 0.1]])


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_dambreak_3d_reduced_against_oracle(hip, oracle, precision):
    geom = cases.dambreak_3d(dp=0.02)
    geom.observer_pos = in_water_probes(geom)
    gpu = cases.DamBreak(hip, geom, precision)
    cpu = cases.DamBreak(oracle, geom, capi.FP64)
    gpu.initialize()
    cpu.initialize()
    fp64 = precision == capi.FP64
    print(f"3D dam break dp 0.02: {gpu.fluid.n} fluid, {gpu.wall.n} wall, precision {precision}", flush=True)
    for k in range(25):
        sort = k == 19
        gpu.time, ng = gpu.advection.exec(gpu.time, sort_now=sort)
        cpu.time, nc = cpu.advection.exec(cpu.time, sort_now=sort)
        if fp64:
            assert ng == nc, (k, ng, nc)
    assert gpu.time == pytest.approx(cpu.time, rel=1e-12 if fp64 else 2e-2)
    gid, cid = gpu.fluid.read("OriginalID"), cpu.fluid.read("OriginalID")
    if fp64:
        assert np.array_equal(gid, cid)
    # FP32-mixed positions differ from FP64 ones by float rounding, so a particle
    # next to a cell face may sort into the neighbouring cell: compare by OriginalID
    go, co = np.argsort(gid), np.argsort(cid)
    assert np.array_equal(gid[go], cid[co])
    tol = {"Position": 1e-12 if fp64 else 1e-4, "Velocity": 1e-9 if fp64 else 3e-2,
           "Compression": 1e-9 if fp64 else 3e-2}
    for name, t in tol.items():
        g, c = gpu.fluid.read(name)[go], cpu.fluid.read(name)[co]
        if name == "Compression":
            g, c = g - 1.0, c - 1.0
        err = rel_l2(g, c)
        print(f"  {name}: rel-L2 {err:.2e} (limit {t:.0e})", flush=True)
        assert err < t, (name, err)
    if fp64:
        # the same relations after the step's rebuild
        for rg, rc in ((gpu.inner, cpu.inner), (gpu.contact, cpu.contact)):
            for a, b in zip(rg.read(), rc.read()):
                assert np.array_equal(a, b)
    pg, eg = gpu.observe()
    pc, ec = cpu.observe()
    print(f"  probes GPU {pg} oracle {pc}", flush=True)
    assert np.all(np.abs(pc[:3]) > 100.0)  # the three inner probes see water
    assert rel_l2(pg, pc) < (1e-9 if fp64 else 2e-2)
    assert eg == pytest.approx(ec, rel=1e-10 if fp64 else 1e-3)
    assert gpu.front.exec() == pytest.approx(cpu.front.exec(), rel=1e-11 if fp64 else 1e-4)
    gpu.close()
    cpu.close()


def test_dambreak_3d_full_size_one_step_against_oracle(hip, oracle):
    geom = cases.dambreak_3d(dp=0.00448)
    nx, ny, nz = (round(geom.meta[k] / geom.dp) for k in ("LL", "LW", "LH"))
    assert nx * ny * nz == len(geom.fluid_pos) == 8_008_376
    print(f"full size: {len(geom.fluid_pos)} fluid, {len(geom.wall_pos)} wall", flush=True)
    cpu = cases.DamBreak(oracle, geom, capi.FP64)
    cpu.initialize()
    ref_lists = [r.read() for r in (cpu.inner, cpu.contact)]
    # closed forms on the initial lattice: all 80 lattice neighbours (|d| < 2.6 dp) for
    # particles two layers or more inside every face, fewer elsewhere
    size = ref_lists[0][0]
    assert size.max() == 80
    assert int(np.count_nonzero(size == 80)) == (nx - 4) * (ny - 4) * (nz - 4)
    h = 1.3 * geom.dp
    adv = cpu.advection_dt.exec()
    assert adv == pytest.approx(0.25 * h / geom.u_ref, rel=1e-15)
    g = float(np.linalg.norm(geom.gravity))
    assert cpu.acoustic_dt.exec() == pytest.approx(0.6 * h / max(geom.c0, (4 * h * g) ** 0.5), rel=1e-15)
    print("  oracle lists and closed forms done", flush=True)
    cpu.time, n_cpu = cpu.advection.exec(cpu.time, sort_now=False)
    ref_state = {k: cpu.fluid.read(k) for k in ("Position", "Velocity", "Compression", "Force")}
    ref_after = [r.read() for r in (cpu.inner, cpu.contact)]
    cpu_time = cpu.time
    cpu.close()
    print(f"  oracle step: {n_cpu} sub-steps, t = {cpu_time:.6e}", flush=True)
    for precision in (capi.FP64, capi.MIXED):
        fp64 = precision == capi.FP64
        gpu = cases.DamBreak(hip, geom, precision)
        gpu.initialize()
        for rel, ref in zip((gpu.inner, gpu.contact), ref_lists):
            for a, b in zip(rel.read(), ref):
                assert np.array_equal(a, b)
        assert gpu.advection_dt.exec() == pytest.approx(adv, rel=1e-15 if fp64 else 1e-6)
        gpu.time, n_gpu = gpu.advection.exec(gpu.time, sort_now=False)
        if fp64:
            assert n_gpu == n_cpu
            assert gpu.time == pytest.approx(cpu_time, rel=1e-13)
        tol = {"Position": 1e-12 if fp64 else 1e-6, "Velocity": 1e-9 if fp64 else 1e-2,
               "Compression": 1e-9 if fp64 else 1e-2, "Force": 1e-9 if fp64 else 2e-2}
        for name, t in tol.items():
            a, b = gpu.fluid.read(name), ref_state[name]
            if name == "Compression":
                a, b = a - 1.0, b - 1.0
            err = rel_l2(a, b)
            print(f"  precision {precision} {name}: rel-L2 {err:.2e} (limit {t:.0e})", flush=True)
            assert err < t, (precision, name, err)
        for rel, ref in zip((gpu.inner, gpu.contact), ref_after):
            for a, b in zip(rel.read(), ref):
                assert np.array_equal(a, b)
        gpu.close()
