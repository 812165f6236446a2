"""Parity of the HIP path (through the C ABI) with the oracle on identical inputs.

Gates: cell lists, Morton sort and neighbour lists bit-exact (both sides use
the canonical ascending orders, so whole arrays are compared); FP64 per-step
forces / compression rates within 1e-10 relative L2; MIXED (float pair math)
reported against the FP64 oracle with the tolerance written in each test."""
import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import cases, ck

pytestmark = pytest.mark.gpu

FP64_GATE = 1e-10


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def jittered_lattice(dim, m, dp, seed, jitter=0.1):
    rng = np.random.default_rng(seed)
    axes = [(np.arange(m) + 0.5) * dp for _ in range(dim)]
    pos = np.stack([g.ravel() for g in np.meshgrid(*axes, indexing="ij")], axis=1)
    return pos + rng.uniform(-jitter * dp, jitter * dp, size=pos.shape)


class Pair:
    """The same fluid (+ optional wall) scene on the HIP library and the oracle."""

    def __init__(self, hip, oracle, fluid, dp, wall=None, normals=None, precision=capi.FP64,
                 rho0=1.0, c0=10.0, bounds=None):
        self.sides = []
        dim = fluid.shape[1]
        lo, up = bounds if bounds is not None else (fluid.min(axis=0) - 0.1, fluid.max(axis=0) + 0.1)
        for lib, prec in ((hip, precision), (oracle, capi.FP64)):
            s = {}
            s["ctx"] = ck.Context(lib)
            s["f"] = ck.Body(s["ctx"], dim, len(fluid), dp, capi.BODY_FLUID, prec, rho0=rho0, c0=c0)
            s["f"].write("Position", fluid)
            s["f"].write("VolumetricMeasure", np.full(len(fluid), dp ** dim))
            s["fc"] = ck.CellLinkedList(s["f"], lo, up)
            s["inner"] = ck.Inner(s["f"], s["fc"])
            s["walls"] = []
            if wall is not None:
                s["w"] = ck.Body(s["ctx"], dim, len(wall), dp, capi.BODY_SOLID, prec, rho0=rho0, c0=c0)
                s["w"].write("Position", wall)
                s["w"].write("VolumetricMeasure", np.full(len(wall), dp ** dim))
                s["w"].write("NormalDirection", normals)
                s["wc"] = ck.CellLinkedList(s["w"], lo, up)
                s["walls"] = [ck.Contact(s["f"], s["w"], s["wc"])]
                ck.UpdateCellLinkedList(s["wc"]).exec()
            ck.UpdateCellLinkedList(s["fc"]).exec()
            ck.UpdateRelation(s["inner"], *s["walls"]).exec()
            self.sides.append(s)
        self.gpu, self.cpu = self.sides

    def both(self, fn):
        return [fn(s) for s in self.sides]

    def write(self, name, value):
        for s in self.sides:
            s["f"].write(name, value)

    def read(self, name):
        return self.gpu["f"].read(name), self.cpu["f"].read(name)

    def close(self):
        for s in self.sides:
            for r in [s["inner"], *s["walls"]]:
                r.close()
            for k in ("fc", "wc", "f", "w"):
                if k in s:
                    s[k].close()
            s["ctx"].close()


@pytest.mark.parametrize("dim", [2, 3])
def test_cell_linked_list_bit_exact(hip, oracle, dim):
    rng = np.random.default_rng(100 + dim)
    pos = rng.uniform(-0.2, 1.2, size=(40000 if dim == 3 else 20000, dim))
    out = []
    for lib in (hip, oracle):
        ctx = ck.Context(lib)
        b = ck.Body(ctx, dim, len(pos), 0.01)
        b.write("Position", pos)
        c = ck.CellLinkedList(b, np.zeros(dim), np.ones(dim))
        ck.UpdateCellLinkedList(c).exec()
        out.append(c.read())
        c.close(), b.close(), ctx.close()
    (go, gi), (co, ci) = out
    assert np.array_equal(go, co)
    assert np.array_equal(gi, ci)


@pytest.mark.parametrize("dim", [2, 3])
def test_cell_linked_list_dense_cells_bit_exact(hip, oracle, dim):
    """Cells of 1 to 150 particles, shuffled so that each cell's particles
    arrive in many runs of the scatter: the per-cell ordering in registers (up
    to 32 particles), in memory above that, and the 32/33 boundary, against the
    oracle."""
    rng = np.random.default_rng(300 + dim)
    dp = 0.01
    probe_ctx = ck.Context(hip)
    probe = ck.Body(probe_ctx, dim, 1, dp)
    mesh = ck.CellLinkedList(probe, np.zeros(dim), np.ones(dim))
    lower, gs, cells = np.array(mesh.lower), mesh.grid_spacing, np.array(mesh.all_cells)
    mesh.close(), probe.close(), probe_ctx.close()
    sizes = [1, 2, 5, 31, 32, 33, 34, 64, 100, 150] * 3
    # distinct cells two apart along every axis, well inside the mesh
    free = np.stack(np.meshgrid(*[np.arange(3, c - 3, 2) for c in cells], indexing="ij"), -1).reshape(-1, dim)
    picked = free[rng.choice(len(free), size=len(sizes), replace=False)]
    pts = [lower + (c + 0.5 + rng.uniform(-0.4, 0.4, size=(k, dim))) * gs for c, k in zip(picked, sizes)]
    pos = np.concatenate(pts)[rng.permutation(sum(sizes))]
    out = []
    for lib in (hip, oracle):
        ctx = ck.Context(lib)
        b = ck.Body(ctx, dim, len(pos), dp)
        b.write("Position", pos)
        c = ck.CellLinkedList(b, np.zeros(dim), np.ones(dim))
        ck.UpdateCellLinkedList(c).exec()
        out.append(c.read())
        c.close(), b.close(), ctx.close()
    (go, gi), (co, ci) = out
    counts = np.diff(go.astype(np.int64))
    assert sorted(counts[counts > 0].tolist()) == sorted(sizes)
    assert np.array_equal(go, co)
    assert np.array_equal(gi, ci)
    for k in np.flatnonzero(counts > 1):
        assert np.all(np.diff(gi[go[k]:go[k + 1]].astype(np.int64)) > 0)


@pytest.mark.parametrize("dim", [2, 3])
def test_particle_sort_bit_exact(hip, oracle, dim):
    rng = np.random.default_rng(200 + dim)
    pos = rng.uniform(0.0, 1.0, size=(30000, dim))
    vel = rng.normal(size=pos.shape)
    res = []
    for lib in (hip, oracle):
        ctx = ck.Context(lib)
        b = ck.Body(ctx, dim, len(pos), 0.005)
        b.write("Position", pos)
        b.write("VolumetricMeasure", np.full(len(pos), 1e-6))
        b.write("Velocity", vel)
        c = ck.CellLinkedList(b, np.zeros(dim), np.ones(dim))
        ck.UpdateCellLinkedList(c).exec()
        ck.ParticleSortCK(b, c).exec()
        res.append([b.read(k) for k in ("SortSequence", "OriginalID", "SortedID", "Position", "Velocity")])
        c.close(), b.close(), ctx.close()
    for g, o in zip(*res):
        assert np.array_equal(g, o)


@pytest.mark.parametrize("dim", [2, 3])
def test_relations_bit_exact(hip, oracle, dim):
    dp = 0.02
    m = 60 if dim == 2 else 22
    fluid = jittered_lattice(dim, m, dp, 7)
    wall = jittered_lattice(dim, m, dp, 8)[::5] - dp * 0.3
    normals = np.tile(np.eye(dim)[-1], (len(wall), 1))
    pr = Pair(hip, oracle, fluid, dp, wall, normals)
    for rg, rc in zip([pr.gpu["inner"], *pr.gpu["walls"]], [pr.cpu["inner"], *pr.cpu["walls"]]):
        for g, o in zip(rg.read(), rc.read()):
            assert np.array_equal(g, o)
    pr.close()


def shuffled(pos, window, seed):
    """Particle order shuffled within consecutive windows (window 0: whole
    array), so that a 256-particle block has neighbours spread over many cells."""
    rng = np.random.default_rng(seed)
    idx = np.arange(len(pos))
    w = window or len(pos)
    for s in range(0, len(pos), w):
        rng.shuffle(idx[s:s + w])
    return pos[idx]


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
@pytest.mark.parametrize("window", [1024, 0])
def test_relations_bit_exact_shuffled(hip, oracle, precision, window):
    """Unsorted particle orders: the staged (MIXED) build meets blocks whose
    distinct-neighbour set exceeds the staging capacity (window 1024) or the
    LDS hash (full shuffle); the lists must stay bit-exact."""
    dp = 0.02
    fluid = shuffled(jittered_lattice(3, 24, dp, 17), window, 5)
    pr = Pair(hip, oracle, fluid, dp, precision=precision)
    for g, o in zip(pr.gpu["inner"].read(), pr.cpu["inner"].read()):
        assert np.array_equal(g, o)
    pr.close()


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_relations_dense_cluster(hip, oracle, precision):
    """A cluster of 4500 particles inside one cell (every pair within the
    cut-off) next to a regular lattice: lists of 4499 entries, beyond the
    tiled build's row code and LDS tile, built by the scan kernel."""
    dp = 0.02
    rng = np.random.default_rng(31)
    lattice = jittered_lattice(2, 40, dp, 30)
    cluster = 0.4 + rng.uniform(0.0, 0.3 * dp, size=(4500, 2))
    fluid = np.concatenate([lattice, cluster])
    pr = Pair(hip, oracle, fluid, dp, precision=precision)
    for g, o in zip(pr.gpu["inner"].read(), pr.cpu["inner"].read()):
        assert np.array_equal(g, o)
    size = pr.gpu["inner"].read()[0]
    assert size.max() >= 4499
    pr.close()


def perturbed_pair(hip, oracle, dim, precision=capi.FP64, seed=1, window=None):
    dp = 0.02
    m = 40 if dim == 2 else 16
    fluid = jittered_lattice(dim, m, dp, seed)
    if window is not None:
        fluid = shuffled(fluid, window, seed + 20)
    # a wall slab below the fluid
    slab = jittered_lattice(dim, m, dp, seed + 1, jitter=0.0)
    wall = slab[slab[:, -1] < 3 * dp].copy()
    wall[:, -1] -= 3 * dp
    normals = np.tile(np.eye(dim)[-1], (len(wall), 1))
    pr = Pair(hip, oracle, fluid, dp, wall, normals, precision, rho0=1000.0, c0=20.0)
    rng = np.random.default_rng(seed + 10)
    n = len(fluid)
    pr.write("Compression", 1.0 + rng.uniform(-0.01, 0.01, n))
    pr.write("CompressionRate", rng.uniform(-0.5, 0.5, n))
    pr.write("Velocity", rng.normal(scale=0.3, size=(n, dim)))
    g = np.zeros(dim)
    g[-1] = -9.81
    for s in pr.sides:
        ck.GravityForceCK(s["f"], g).exec()
    return pr


@pytest.mark.parametrize("dim", [2, 3])
def test_summation_regularization_setup_fp64(hip, oracle, dim):
    pr = perturbed_pair(hip, oracle, dim)
    pr.both(lambda s: ck.CompressionSummation(s["inner"], *s["walls"]).exec())
    g, c = pr.read("CompressionSummation")
    assert rel_l2(g, c) < FP64_GATE
    pr.both(lambda s: ck.DensityRegularization(s["f"]).exec())
    pr.both(lambda s: ck.AdvectionStepSetup(s["f"]).exec())
    for name in ("Compression", "Density", "VolumetricMeasure"):
        g, c = pr.read(name)
        assert rel_l2(g, c) < FP64_GATE, name
    pr.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("riemann", [capi.RIEMANN_ACOUSTIC, capi.RIEMANN_NONE, capi.RIEMANN_DISSIPATIVE])
def test_acoustic_step_fp64(hip, oracle, dim, riemann):
    pr = perturbed_pair(hip, oracle, dim)
    dt = 2e-5
    pr.both(lambda s: ck.AcousticStep1stHalf(s["inner"], *s["walls"], riemann=riemann).exec(dt))
    for name in ("Pressure", "Force", "CompressionRate", "Velocity", "Displacement"):
        g, c = pr.read(name)
        assert rel_l2(g, c) < FP64_GATE, name
    pr.both(lambda s: ck.AcousticStep2ndHalf(s["inner"], *s["walls"], riemann=riemann).exec(dt))
    for name in ("Force", "CompressionRate", "Compression", "Density", "Displacement"):
        g, c = pr.read(name)
        assert rel_l2(g, c) < FP64_GATE, name
    dts = pr.both(lambda s: ck.AcousticTimeStepCK(s["f"]).exec())
    assert dts[0] == pytest.approx(dts[1], rel=1e-13)
    dts = pr.both(lambda s: ck.AdvectionTimeStepCK(s["f"], 1.0).exec())
    assert dts[0] == pytest.approx(dts[1], rel=1e-13)
    g = np.zeros(dim)
    g[-1] = -9.81
    es = pr.both(lambda s: ck.TotalMechanicalEnergyCK(s["f"], g).exec())
    assert es[0] == pytest.approx(es[1], rel=1e-12)
    fr = pr.both(lambda s: ck.UpperFrontInAxisDirectionCK(s["f"], 0).exec())
    assert fr[0] == fr[1]
    pr.close()


@pytest.mark.parametrize("dim,window", [(2, None), (3, None), (3, 1024), (3, 0)])
def test_acoustic_step_mixed_against_fp64_oracle(hip, oracle, dim, window):
    """FP32 pair interactions: per-step force and compression rate within 2e-4
    relative L2 of the FP64 oracle (float rounding of positions at ~1e-7 of the
    domain and of the pair sums).  Shuffled orders send blocks through the
    unstaged (global gather) path of the passes."""
    pr = perturbed_pair(hip, oracle, dim, precision=capi.MIXED, window=window)
    dt = 2e-5
    pr.both(lambda s: ck.CompressionSummation(s["inner"], *s["walls"]).exec())
    g, c = pr.read("CompressionSummation")
    assert rel_l2(g, c) < 1e-6
    pr.both(lambda s: ck.AcousticStep1stHalf(s["inner"], *s["walls"]).exec(dt))
    for name, tol in (("Pressure", 1e-5), ("Force", 2e-4), ("CompressionRate", 2e-4), ("Velocity", 1e-5)):
        g, c = pr.read(name)
        assert rel_l2(g, c) < tol, (name, rel_l2(g, c))
    pr.both(lambda s: ck.AcousticStep2ndHalf(s["inner"], *s["walls"]).exec(dt))
    for name, tol in (("Force", 2e-4), ("CompressionRate", 2e-4), ("Compression", 1e-6)):
        g, c = pr.read(name)
        assert rel_l2(g, c) < tol, (name, rel_l2(g, c))
    pr.close()


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_dambreak_2d_short_run(hip, oracle, precision):
    """25 advection steps of the 2D dam break through the single-call driver,
    with a sort after the 20th, against the oracle.  FP64: same sub-step count,
    state within 1e-9 relative L2.  MIXED: state within the float tolerances
    written below."""
    geom = cases.dambreak_2d(dp=0.025)
    gpu = cases.DamBreak(hip, geom, precision)
    cpu = cases.DamBreak(oracle, geom, capi.FP64)
    gpu.initialize()
    cpu.initialize()
    for k in range(25):
        sort = k == 19
        gpu.time, ng = gpu.advection.exec(gpu.time, sort_now=sort)
        cpu.time, nc = cpu.advection.exec(cpu.time, sort_now=sort)
        if precision == capi.FP64:
            assert ng == nc
    fp64 = precision == capi.FP64
    assert gpu.time == pytest.approx(cpu.time, rel=1e-12 if fp64 else 2e-2)
    assert np.array_equal(gpu.fluid.read("OriginalID"), cpu.fluid.read("OriginalID"))
    tol = {"Position": 1e-12 if fp64 else 1e-4, "Velocity": 1e-9 if fp64 else 3e-2,
           "Compression": 1e-9 if fp64 else 3e-2}
    for name, t in tol.items():
        g, c = gpu.fluid.read(name), cpu.fluid.read(name)
        if name == "Compression":
            g, c = g - 1.0, c - 1.0
        assert rel_l2(g, c) < t, (name, rel_l2(g, c))
    pg, eg = gpu.observe()
    pc, ec = cpu.observe()
    assert eg == pytest.approx(ec, rel=1e-10 if fp64 else 1e-3)
    # the wall probe (no water there yet at t ~ 0.2: both zero) ...
    assert np.allclose(pg, pc, rtol=1e-9 if fp64 else 2e-2, atol=1e-12)
    # ... and Interpolation<Contact<Real>> at points inside the water column, on
    # the same states: pressure, density and velocity
    pts = np.array([[0.3, 0.2], [1.0, 0.5], [1.7, 0.1]])
    probes = {}
    for key, sc in (("gpu", gpu), ("cpu", cpu)):
        obs = ck.Body(sc.ctx, 2, len(pts), geom.dp, capi.BODY_OBSERVER, sc.precision, rho0=geom.rho0, c0=geom.c0)
        obs.write("Position", pts)
        obs.write("VolumetricMeasure", np.full(len(pts), geom.dp ** 2))
        rel = ck.Contact(obs, sc.fluid, sc.fluid_cll)
        ck.UpdateRelation(rel).exec()
        probes[key] = [ck.InterpolationContact(rel, f).exec() for f in ("Pressure", "Density", "Velocity")]
        rel.close()
        obs.close()
    assert np.all(np.abs(probes["cpu"][0]) > 1e-3)  # the points see water
    for g, c, tol in zip(probes["gpu"], probes["cpu"], ((1e-9, 2e-2), (1e-12, 1e-5), (1e-9, 3e-2))):
        assert rel_l2(g, c) < tol[0 if fp64 else 1], (rel_l2(g, c), g, c)
    gpu.close()
    cpu.close()


def test_density_sum_microbench_reduced(hip, oracle):
    """The neighbour-search + density-summation microbenchmark at 48^3 (the full
    16M case is timed in bench.py): cell list, neighbour lists and counts
    bit-exact, FP64 density sum within 1e-12, MIXED within 1e-6."""
    m, dp = 48, 1.0 / 48
    rng = np.random.default_rng(20261202)
    axes = [(np.arange(m) + 0.5) * dp for _ in range(3)]
    pos = np.stack([g.ravel() for g in np.meshgrid(*axes, indexing="ij")], axis=1)
    pos += rng.uniform(-0.1 * dp, 0.1 * dp, size=pos.shape)
    res = {}
    for key, lib, prec in (("gpu64", hip, capi.FP64), ("gpu32", hip, capi.MIXED), ("cpu", oracle, capi.FP64)):
        ctx = ck.Context(lib)
        b = ck.Body(ctx, 3, len(pos), dp, capi.BODY_FLUID, prec)
        b.write("Position", pos)
        b.write("VolumetricMeasure", np.full(len(pos), dp ** 3))
        lo, up = np.zeros(3) - 4 * dp, np.ones(3) + 4 * dp
        c = ck.CellLinkedList(b, lo, up)
        inner = ck.Inner(b, c)
        ck.UpdateCellLinkedList(c).exec()
        ck.UpdateRelation(inner).exec()
        ck.CompressionSummation(inner).exec()
        res[key] = (c.read(), inner.read(), b.read("CompressionSummation"))
        inner.close(), c.close(), b.close(), ctx.close()
    for key in ("gpu64", "gpu32"):
        (go, gi), rel_g, _ = res[key]
        (co, ci), rel_c, _ = res["cpu"]
        assert np.array_equal(go, co) and np.array_equal(gi, ci)
        for g, o in zip(rel_g, rel_c):
            assert np.array_equal(g, o)
    assert rel_l2(res["gpu64"][2], res["cpu"][2]) < 1e-12
    assert rel_l2(res["gpu32"][2], res["cpu"][2]) < 1e-6
