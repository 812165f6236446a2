"""Pin the CPU oracle against known answers, closed forms, invariants and brute
force (it is the checker of the HIP path, so it is checked first).

Every check is independent of the oracle's own code: Wendland C2 is evaluated
from its closed form, cells and Morton codes by numpy arithmetic, neighbour
sets by an O(N^2) search, pair forces from the hand-derived two-particle
formulas of acoustic_step_{1st,2nd}_half.hpp."""
import ctypes as C
import math

import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import ck


def wendland_w(q):  # kernel_wendland_c2.cpp:17-20
    return (1.0 - 0.5 * q) ** 4 * (1.0 + 2.0 * q)


def wendland_dw(q):  # kernel_wendland_c2.cpp:32-35
    return 0.625 * (q - 2.0) ** 3 * q


def dimfac(d):
    return 7.0 / (4.0 * math.pi) if d == 2 else 21.0 / (16.0 * math.pi)


def make_body(lib, pos, dp=0.1, kind=capi.BODY_FLUID, rho0=1.0, c0=10.0, precision=capi.FP64):
    pos = np.asarray(pos, dtype=np.float64)
    ctx = ck.Context(lib)
    b = ck.Body(ctx, pos.shape[1], len(pos), dp, kind, precision, rho0=rho0, c0=c0)
    b.write("Position", pos)
    b.write("VolumetricMeasure", np.full(len(pos), dp ** pos.shape[1]))
    return ctx, b


def domain(pos, pad=1.0):
    pos = np.asarray(pos)
    return pos.min(axis=0) - pad, pos.max(axis=0) + pad


# ---------------------------------------------------------------- scan
def test_exclusive_scan_known_answer(oracle):
    # tests/unit_tests_src/shared/particle_dynamics/configuration_dynamics/test_exclusive_scan.cpp:8-9
    data = np.array([3, 2, 3, 5, 0, 1, 3, 2, 5, 1, 0], dtype=np.uint32)
    out = np.zeros_like(data)
    f = oracle.lib.sphx_oracle_exclusive_scan
    f.restype = C.c_uint32
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    total = f(data.ctypes.data, out.ctypes.data, data.size)
    assert out.tolist() == [0, 3, 5, 8, 13, 13, 14, 17, 19, 24, 25]
    assert total == 25


# ---------------------------------------------------------------- kernel
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("node", [0, 3, 7, 10, 14, 19])
def test_kernel_value_and_gradient_at_table_nodes(oracle, dim, node):
    """At q on a table node the 4-point interpolation returns the node value:
    W and dW must equal the Wendland closed form there."""
    dp = 0.1
    h = 1.3 * dp
    q = node * 0.1
    r = q * h
    pos = np.zeros((2, dim))
    pos[1, 0] = r
    ctx, b = make_body(oracle, pos, dp)
    lo, up = domain(pos)
    cll = ck.CellLinkedList(b, lo, up)
    inner = ck.Inner(b, cll)
    ck.UpdateCellLinkedList(cll).exec()
    ck.UpdateRelation(inner).exec()
    V = dp ** dim
    ck.CompressionSummation(inner).exec()
    S = b.read("CompressionSummation")
    W0 = dimfac(dim) / h ** dim * wendland_w(0.0)
    W = dimfac(dim) / h ** dim * wendland_w(q)
    assert S[0] == pytest.approx((W0 + W) * V, rel=1e-12)
    # pass 1 with only inner pressure: F_0 = -(p0 + p1) dW V0 V1 e_01
    p = np.array([2.0, 5.0])
    b.write("Pressure", p)
    b.write("Compression", 1.0 + p / (1.0 * 10.0 ** 2))  # keeps p after the initialize
    b.write("Mass", np.full(2, 1.0))
    ck.AcousticStep1stHalf(inner, riemann=capi.RIEMANN_NONE).exec(0.0)
    F = b.read("Force")
    dW = dimfac(dim) / h ** (dim + 1) * wendland_dw(q)
    if node == 0:
        assert np.allclose(F, 0.0)
    else:
        e01 = -np.eye(dim)[0]  # x_0 - x_1 points along -x
        expect = -(p[0] + p[1]) * dW * V * V * e01
        assert np.allclose(F[0], expect, rtol=1e-10, atol=0)
        assert np.allclose(F[1], -expect, rtol=1e-10, atol=0)
    ctx.close()


@pytest.mark.parametrize("dim,closed", [(2, None), (3, None)])
def test_lattice_number_density_matches_closed_form(oracle, dim, closed):
    """Interior lattice sum of W V (CompressionSummation on a perfect lattice)
    against the closed-form Wendland sum; differs only by the tabulation error."""
    dp = 0.1
    n = 9 if dim == 3 else 15
    axes = [np.arange(n) * dp for _ in range(dim)]
    pos = np.stack([g.ravel() for g in np.meshgrid(*axes, indexing="ij")], axis=1)
    ctx, b = make_body(oracle, pos, dp)
    lo, up = domain(pos)
    cll = ck.CellLinkedList(b, lo, up)
    inner = ck.Inner(b, cll)
    ck.UpdateCellLinkedList(cll).exec()
    ck.UpdateRelation(inner).exec()
    ck.CompressionSummation(inner).exec()
    S = b.read("CompressionSummation")
    centre = np.argmin(np.sum((pos - pos.mean(axis=0)) ** 2, axis=1))
    h = 1.3
    s = 0.0
    rng = range(-3, 4)
    for idx in np.ndindex(*([7] * dim)):
        q = math.sqrt(sum((k - 3) ** 2 for k in idx)) / h
        if q < 2.0:
            s += wendland_w(q)
    expect = s * dimfac(dim) / h ** dim
    assert S[centre] == pytest.approx(expect, rel=5e-5)
    ctx.close()


# ---------------------------------------------------------------- mesh, CLL, Morton
def np_cells(pos, lower, gs, grid_points):
    c = np.floor((pos - lower) / gs).astype(np.int64)
    return np.clip(c, 0, np.asarray(grid_points) - 2)


@pytest.mark.parametrize("dim", [2, 3])
def test_cell_linked_list_against_closed_form(oracle, dim):
    rng = np.random.default_rng(7)
    n = 3000
    pos = rng.uniform(-0.3, 1.3, size=(n, dim))
    lo, up = np.zeros(dim), np.ones(dim)
    ctx, b = make_body(oracle, pos, dp=0.05)
    cll = ck.CellLinkedList(b, lo, up)  # particles outside the bounds are clamped
    ck.UpdateCellLinkedList(cll).exec()
    off, idx = cll.read()
    gs = cll.grid_spacing
    assert gs == 2.0 * 1.3 * 0.05
    # Mesh(bounds, gs, 2): lower = lo - 2 gs, cells = ceil((up + 2gs - lower)/gs)
    lower = lo - 2 * gs
    cells = np.ceil((up + 2 * gs - lower) / gs).astype(int)
    assert tuple(cells) == cll.all_cells
    c = np_cells(pos, lower, gs, cells + 1)
    lin = c[:, 0] * cells[1] + c[:, 1] if dim == 2 else (c[:, 0] * cells[1] + c[:, 1]) * cells[2] + c[:, 2]
    counts = np.bincount(lin, minlength=int(np.prod(cells)))
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32))
    for cell in np.unique(lin)[:200]:
        members = np.sort(np.nonzero(lin == cell)[0])
        assert np.array_equal(idx[off[cell]:off[cell + 1]], members)
    ctx.close()


def morton_np(c):
    # Mesh::MortonCode spreads the 10 low bits of a cell index to every third
    # bit (mask 0x9249249) in 2D and 3D alike; axis a is shifted left by a
    key = np.zeros(len(c), dtype=np.uint64)
    for bit in range(10):
        for axis in range(c.shape[1]):
            key |= ((c[:, axis].astype(np.uint64) >> bit) & 1) << (bit * 3 + axis)
    return key.astype(np.uint32)


@pytest.mark.parametrize("dim", [2, 3])
def test_particle_sort_morton_keys_and_permutation(oracle, dim):
    rng = np.random.default_rng(11)
    n = 2500
    pos = rng.uniform(0.0, 1.0, size=(n, dim))
    vel = rng.normal(size=(n, dim))
    ctx, b = make_body(oracle, pos, dp=0.02)
    b.write("Velocity", vel)
    lo, up = np.zeros(dim), np.ones(dim)
    cll = ck.CellLinkedList(b, lo, up)
    ck.UpdateCellLinkedList(cll).exec()
    ck.ParticleSortCK(b, cll).exec()
    seq = b.read("SortSequence")
    oid = b.read("OriginalID")
    sid = b.read("SortedID")
    gs = cll.grid_spacing
    c = np_cells(pos, lo - 2 * gs, gs, np.array(cll.all_cells) + 1)
    keys = morton_np(c)
    assert np.all(np.diff(seq.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(oid), np.arange(n))
    assert np.array_equal(seq, keys[oid])
    assert np.array_equal(sid[oid], np.arange(n))
    assert np.array_equal(b.read("Position"), pos[oid])
    assert np.array_equal(b.read("Velocity"), vel[oid])
    # stable: equal keys keep ascending original index
    same = seq[1:] == seq[:-1]
    assert np.all(oid[1:][same] > oid[:-1][same])
    ctx.close()


# ---------------------------------------------------------------- relations
def brute_force_neighbors(src, tar, inv_h, inner):
    out = []
    for i in range(len(src)):
        s = inv_h * (src[i] - tar)
        r2 = s[:, 0] * s[:, 0] + s[:, 1] * s[:, 1]
        if src.shape[1] == 3:
            r2 = r2 + s[:, 2] * s[:, 2]
        m = r2 < 4.0
        if inner:
            m[i] = False
        out.append(set(np.nonzero(m)[0].tolist()))
    return out


@pytest.mark.parametrize("dim", [2, 3])
def test_inner_and_contact_relations_against_brute_force(oracle, dim):
    rng = np.random.default_rng(3)
    dp = 0.05
    m = 14 if dim == 2 else 8
    axes = [np.arange(m) * dp for _ in range(dim)]
    lat = np.stack([g.ravel() for g in np.meshgrid(*axes, indexing="ij")], axis=1)
    fluid = lat + rng.uniform(-0.2 * dp, 0.2 * dp, size=lat.shape)
    wall = lat[:: 3] + np.array([0.01] * dim)
    ctx = ck.Context(oracle)
    fb = ck.Body(ctx, dim, len(fluid), dp)
    wb = ck.Body(ctx, dim, len(wall), dp, capi.BODY_SOLID)
    fb.write("Position", fluid)
    wb.write("Position", wall)
    lo, up = domain(np.vstack([fluid, wall]), 0.1)
    fc, wc = ck.CellLinkedList(fb, lo, up), ck.CellLinkedList(wb, lo, up)
    inner, contact = ck.Inner(fb, fc), ck.Contact(fb, wb, wc)
    for c in (fc, wc):
        ck.UpdateCellLinkedList(c).exec()
    ck.UpdateRelation(inner, contact).exec()
    inv_h = 1.0 / (1.3 * dp)
    for rel, tar, is_inner in ((inner, fluid, True), (contact, wall, False)):
        size, off, idx = rel.read()
        expect = brute_force_neighbors(fluid, tar, inv_h, is_inner)
        assert np.array_equal(np.diff(off), size)
        for i in range(len(fluid)):
            got = idx[off[i]:off[i + 1]]
            assert len(set(got.tolist())) == len(got)
            assert set(got.tolist()) == expect[i]
    # inner symmetry: j in N(i) <=> i in N(j)
    size, off, idx = inner.read()
    pairs = {(i, int(j)) for i in range(len(fluid)) for j in idx[off[i]:off[i + 1]]}
    assert all((j, i) in pairs for (i, j) in pairs)
    ctx.close()


@pytest.mark.parametrize("dim,expected", [(2, 20), (3, 80)])
def test_interior_lattice_neighbor_count(oracle, dim, expected):
    dp = 0.1
    n = 9
    axes = [np.arange(n) * dp for _ in range(dim)]
    pos = np.stack([g.ravel() for g in np.meshgrid(*axes, indexing="ij")], axis=1)
    ctx, b = make_body(oracle, pos, dp)
    lo, up = domain(pos)
    cll = ck.CellLinkedList(b, lo, up)
    inner = ck.Inner(b, cll)
    ck.UpdateCellLinkedList(cll).exec()
    ck.UpdateRelation(inner).exec()
    size, _, _ = inner.read()
    centre = np.argmin(np.sum((pos - pos.mean(axis=0)) ** 2, axis=1))
    assert size[centre] == expected
    ctx.close()


# ---------------------------------------------------------------- pair physics
def two_particle_scene(lib, dim, q_node, rho0=1.0, c0=10.0):
    dp = 0.1
    h = 1.3 * dp
    pos = np.zeros((2, dim))
    pos[1, 0] = q_node * 0.1 * h
    ctx, b = make_body(lib, pos, dp, rho0=rho0, c0=c0)
    lo, up = domain(pos)
    cll = ck.CellLinkedList(b, lo, up)
    inner = ck.Inner(b, cll)
    ck.UpdateCellLinkedList(cll).exec()
    ck.UpdateRelation(inner).exec()
    dW = dimfac(dim) / h ** (dim + 1) * wendland_dw(q_node * 0.1)
    return ctx, b, inner, dW, dp ** dim


@pytest.mark.parametrize("dim", [2, 3])
def test_first_half_continuity_and_velocity_update(oracle, dim):
    """Ċ_i = C_i (p_i - p_j)/(rho0 c0) dW V_j; v += (Fp + F)/m dt."""
    rho0, c0 = 1.0, 10.0
    ctx, b, inner, dW, V = two_particle_scene(oracle, dim, 8, rho0, c0)
    C0 = np.array([1.01, 1.002])
    b.write("Compression", C0)
    b.write("Mass", np.full(2, 0.5))
    b.write("ForcePrior", np.tile(np.eye(dim)[-1] * -0.5, (2, 1)))
    dt = 1e-3
    ck.AcousticStep1stHalf(inner, riemann=capi.RIEMANN_ACOUSTIC).exec(dt)
    C = b.read("Compression")
    assert np.allclose(C, C0)  # CompressionRate was 0
    p = rho0 * c0 ** 2 * (C0 - 1.0)
    assert np.allclose(b.read("Pressure"), p, rtol=1e-12)
    dC = b.read("CompressionRate")
    assert dC[0] == pytest.approx(C0[0] * (p[0] - p[1]) / (rho0 * c0) * dW * V, rel=1e-10)
    assert dC[1] == pytest.approx(C0[1] * (p[1] - p[0]) / (rho0 * c0) * dW * V, rel=1e-10)
    F = b.read("Force")
    e = np.zeros(dim)
    e[0] = -1.0
    assert np.allclose(F[0], -(p[0] + p[1]) * dW * V * V * e, rtol=1e-10)
    v = b.read("Velocity")
    g = np.eye(dim)[-1] * -0.5
    assert np.allclose(v[0], (g + F[0]) / 0.5 * dt, rtol=1e-10)
    ctx.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_second_half_dissipation_and_divergence(oracle, dim):
    """F_i = V_i rho0c0 u min(3 max(u,0)/c0, 1) dW V_j e_ij, u = (v_i - v_j).e_ij;
    Ċ_i += C_i u dW V_j; C += dt/2 Ċ."""
    rho0, c0 = 1.0, 10.0
    ctx, b, inner, dW, V = two_particle_scene(oracle, dim, 12, rho0, c0)
    vel = np.zeros((2, dim))
    vel[0, 0] = -2.0  # particle 0 moves toward particle 1: u = (v0 - v1).e01 > 0
    vel[1, 0] = 0.5
    b.write("Velocity", vel)
    dt = 1e-3
    ck.AcousticStep2ndHalf(inner, riemann=capi.RIEMANN_ACOUSTIC).exec(dt)
    e01 = np.zeros(dim)
    e01[0] = -1.0
    u = np.dot(vel[0] - vel[1], e01)
    assert u > 0
    pj = rho0 * c0 * u * min(3.0 * max(u, 0.0) / c0, 1.0)
    F = b.read("Force")
    assert np.allclose(F[0], V * pj * dW * V * e01, rtol=1e-10)
    dC = b.read("CompressionRate")
    assert dC[0] == pytest.approx(1.0 * u * dW * V, rel=1e-10)
    assert b.read("Compression")[0] == pytest.approx(1.0 + 0.5 * dt * dC[0], rel=1e-14)
    dpos = b.read("Displacement")
    assert np.allclose(dpos, vel * dt * 0.5, rtol=1e-14)
    ctx.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_wall_pressure_and_dissipation(oracle, dim):
    """p_w = p_i + rho_i r max(0, (Fp_i/m_i - a_w).(-e)); F_i -= (p_i + p_w) dW V_w e V_i;
    2nd half: n' = sgn(e.n) n, u = 2(v_i - v_w).n'."""
    rho0, c0 = 1.0, 10.0
    dp = 0.1
    h = 1.3 * dp
    q = 0.9
    fpos = np.zeros((1, dim))
    fpos[0, -1] = q * h  # fluid above the wall particle
    wpos = np.zeros((1, dim))
    ctx = ck.Context(oracle)
    fb = ck.Body(ctx, dim, 1, dp, rho0=rho0, c0=c0)
    wb = ck.Body(ctx, dim, 1, dp, capi.BODY_SOLID)
    fb.write("Position", fpos)
    fb.write("VolumetricMeasure", np.full(1, dp ** dim))
    wb.write("Position", wpos)
    wb.write("VolumetricMeasure", np.full(1, dp ** dim))
    n = np.zeros((1, dim))
    n[0, -1] = 1.0
    wb.write("NormalDirection", n)
    lo, up = domain(np.vstack([fpos, wpos]))
    fc, wc = ck.CellLinkedList(fb, lo, up), ck.CellLinkedList(wb, lo, up)
    inner, contact = ck.Inner(fb, fc), ck.Contact(fb, wb, wc)
    for c in (fc, wc):
        ck.UpdateCellLinkedList(c).exec()
    ck.UpdateRelation(inner, contact).exec()
    m = 0.7
    fb.write("Mass", np.full(1, m))
    g = np.zeros((1, dim))
    g[0, -1] = -1.0 * m
    fb.write("ForcePrior", g)
    C0 = 1.003
    fb.write("Compression", np.full(1, C0))
    ck.AcousticStep1stHalf(inner, contact).exec(0.0)
    V = dp ** dim
    dW = dimfac(dim) / h ** (dim + 1) * wendland_dw(q)
    p_i = rho0 * c0 ** 2 * (C0 - 1.0)
    rho_i = C0 * rho0
    e = np.zeros(dim)
    e[-1] = 1.0
    face = np.dot(g[0] / m - 0.0, -e)
    p_w = p_i + rho_i * q * h * max(0.0, face)
    F = fb.read("Force")
    assert np.allclose(F[0], -(p_i + p_w) * dW * V * e * V, rtol=1e-10)
    dC = fb.read("CompressionRate")
    assert dC[0] == pytest.approx(C0 * (p_i - p_w) / (rho0 * c0) * dW * V, rel=1e-10)
    # 2nd half with the fluid moving into the wall
    vel = np.zeros((1, dim))
    vel[0, -1] = -0.8
    fb.write("Velocity", vel)
    fb.write("Force", np.zeros((1, dim)))
    fb.write("CompressionRate", np.zeros(1))
    ck.AcousticStep2ndHalf(inner, contact).exec(0.0)
    vd = 2.0 * (vel[0] - 0.0)
    nprime = np.sign(np.dot(e, n[0])) * n[0]
    u = np.dot(vd, nprime)
    pj = rho0 * c0 * u * min(3.0 * max(u, 0.0) / c0, 1.0)
    assert np.allclose(fb.read("Force")[0], V * pj * dW * V * nprime, rtol=1e-10, atol=1e-300)
    assert fb.read("CompressionRate")[0] == pytest.approx(C0 * np.dot(vd, e) * dW * V, rel=1e-10)
    ctx.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_inner_pressure_force_conserves_momentum(oracle, dim):
    rng = np.random.default_rng(5)
    dp = 0.05
    m = 12 if dim == 2 else 7
    axes = [np.arange(m) * dp for _ in range(dim)]
    pos = np.stack([g.ravel() for g in np.meshgrid(*axes, indexing="ij")], axis=1)
    pos += rng.uniform(-0.15 * dp, 0.15 * dp, size=pos.shape)
    ctx, b = make_body(oracle, pos, dp)
    lo, up = domain(pos)
    cll = ck.CellLinkedList(b, lo, up)
    inner = ck.Inner(b, cll)
    ck.UpdateCellLinkedList(cll).exec()
    ck.UpdateRelation(inner).exec()
    b.write("Compression", 1.0 + rng.uniform(0, 0.02, len(pos)))
    b.write("Mass", np.full(len(pos), 1.0))
    ck.AcousticStep1stHalf(inner).exec(0.0)
    F = b.read("Force")
    assert np.abs(F.sum(axis=0)).max() < 1e-12 * np.abs(F).sum()
    # dissipation of the 2nd half is antisymmetric too
    b.write("Velocity", rng.normal(size=pos.shape))
    ck.AcousticStep2ndHalf(inner).exec(0.0)
    F = b.read("Force")
    assert np.abs(F.sum(axis=0)).max() < 1e-12 * np.abs(F).sum()
    ctx.close()


# ---------------------------------------------------------------- reductions
@pytest.mark.parametrize("dim", [2, 3])
def test_time_steps_and_energy_closed_form(oracle, dim):
    rng = np.random.default_rng(9)
    n = 500
    pos = rng.uniform(0, 1, size=(n, dim))
    ctx, b = make_body(oracle, pos, dp=0.02, rho0=2.0, c0=15.0)
    vel = rng.normal(size=(n, dim))
    F = rng.normal(size=(n, dim)) * 1e-3
    Fp = rng.normal(size=(n, dim)) * 1e-3
    m = rng.uniform(0.5, 1.5, n) * 1e-3
    for k, v in (("Velocity", vel), ("Force", F), ("ForcePrior", Fp), ("Mass", m)):
        b.write(k, v)
    h = 1.3 * 0.02
    sig = np.maximum(15.0 + np.linalg.norm(vel, axis=1),
                     np.sqrt(4.0 * h * np.linalg.norm(F + Fp, axis=1) / m))
    assert ck.AcousticTimeStepCK(b).exec() == pytest.approx(0.6 * h / sig.max(), rel=1e-14)
    vmax = np.sqrt((vel ** 2).sum(axis=1).max())
    assert ck.AdvectionTimeStepCK(b, 1.0).exec() == pytest.approx(0.25 * h / max(vmax, 1.0), rel=1e-14)
    assert ck.AdvectionTimeStepCK(b, 100.0).exec() == pytest.approx(0.25 * h / 100.0, rel=1e-14)
    g = np.zeros(dim)
    g[-1] = -9.81
    e = np.sum(0.5 * m * (vel ** 2).sum(axis=1) + m * (pos @ -g))
    assert ck.TotalMechanicalEnergyCK(b, g).exec() == pytest.approx(e, rel=1e-12)
    assert ck.UpperFrontInAxisDirectionCK(b, 0).exec() == pos[:, 0].max()
    ctx.close()


# ---------------------------------------------------------------- streaming updates
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("flow", [capi.FLOW_FREE_SURFACE, capi.FLOW_INTERNAL])
def test_summation_regularization_setup_position_closed_form(oracle, dim, flow):
    """Isolated particles: S = W0 Vref (W0 = dimfac / h^d, as Wendland(0) = 1);
    a dense interior lattice: S > 1.  FreeSurface: C = max(S, 1); Internal:
    C = S; rho = C rho0 (density_regularization.hpp:85-118).  Setup: V = m/rho,
    dpos = 0; position update: x += dpos (fluid_time_step_ck.h:133-184).
    Gravity: F_prior = m g (force_prior_ck.hpp:37-43)."""
    dp, rho0 = 0.1, 1000.0
    h = 1.3 * dp
    m = 7 if dim == 2 else 5
    axes = [np.arange(m) * dp for _ in range(dim)]
    lattice = np.stack([g.ravel() for g in np.meshgrid(*axes, indexing="ij")], axis=1)
    lone = np.full((2, dim), 10.0)
    lone[1, 0] += 1.0
    pos = np.vstack([lattice, lone])
    n = len(pos)
    ctx, b = make_body(oracle, pos, dp, rho0=rho0)
    lo, up = domain(pos)
    cll = ck.CellLinkedList(b, lo, up)
    inner = ck.Inner(b, cll)
    ck.UpdateCellLinkedList(cll).exec()
    ck.UpdateRelation(inner).exec()
    mass = np.linspace(0.9, 1.1, n)
    b.write("Mass", mass)
    g = np.zeros(dim)
    g[-1] = -9.81
    ck.GravityForceCK(b, g).exec()
    assert np.allclose(b.read("ForcePrior"), mass[:, None] * g[None, :], rtol=1e-15)
    ck.CompressionSummation(inner).exec()
    S = b.read("CompressionSummation")
    V = dp ** dim
    assert S[-1] == pytest.approx(dimfac(dim) / h ** dim * V, rel=1e-14)
    assert S[-2] == pytest.approx(S[-1], rel=1e-15)
    centre = np.argmin(np.sum((lattice - lattice.mean(axis=0)) ** 2, axis=1))
    assert S[centre] > 1.0 > S[-1]
    ck.DensityRegularization(b, flow).exec()
    C = b.read("Compression")
    expect = np.maximum(S, 1.0) if flow == capi.FLOW_FREE_SURFACE else S
    assert np.array_equal(C, expect)
    assert np.allclose(b.read("Density"), expect * rho0, rtol=1e-15)
    b.write("Displacement", np.ones((n, dim)))
    ck.AdvectionStepSetup(b).exec()
    assert np.allclose(b.read("VolumetricMeasure"), mass / (expect * rho0), rtol=1e-15)
    assert not b.read("Displacement").any()
    rng = np.random.default_rng(3)
    dpos = rng.normal(scale=1e-3, size=(n, dim))
    b.write("Displacement", dpos)
    ck.UpdateParticlePosition(b).exec()
    assert np.array_equal(b.read("Position"), pos + dpos)
    ctx.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_riemann_variants_two_particle(oracle, dim):
    """NoRiemannSolverCK: no dissipation (F of the 2nd half = 0, no pressure
    jump in the 1st-half continuity); the plain dissipative solver has no
    limiter: PJump = rho0 c0 u (riemann_solver_ck.h:46-173)."""
    rho0, c0 = 1.0, 10.0
    vel = np.zeros((2, dim))
    vel[0, 0] = -0.5  # u = 0.5 + 0.2 > 0 with 3u/c0 = 0.21 < 1: the limiter matters
    vel[1, 0] = 0.2
    e01 = np.zeros(dim)
    e01[0] = -1.0
    u = np.dot(vel[0] - vel[1], e01)
    out = {}
    for kind in (capi.RIEMANN_ACOUSTIC, capi.RIEMANN_NONE, capi.RIEMANN_DISSIPATIVE):
        ctx, b, inner, dW, V = two_particle_scene(oracle, dim, 11, rho0, c0)
        b.write("Compression", np.array([1.01, 1.0]))
        ck.AcousticStep1stHalf(inner, riemann=kind).exec(0.0)
        dC1 = b.read("CompressionRate")[0]
        b.write("Velocity", vel)
        b.write("CompressionRate", np.zeros(2))
        ck.AcousticStep2ndHalf(inner, riemann=kind).exec(0.0)
        out[kind] = (dC1, b.read("Force")[0], dW, V)
        ctx.close()
    dC1, F, dW, V = out[capi.RIEMANN_NONE]
    assert dC1 == 0.0 and not F.any()
    p0 = rho0 * c0 ** 2 * 0.01
    for kind in (capi.RIEMANN_ACOUSTIC, capi.RIEMANN_DISSIPATIVE):
        assert out[kind][0] == pytest.approx(1.01 * p0 / (rho0 * c0) * dW * V, rel=1e-10)
    limiter = min(3.0 * max(u, 0.0) / c0, 1.0)
    assert np.allclose(out[capi.RIEMANN_ACOUSTIC][1], V * rho0 * c0 * u * limiter * dW * V * e01, rtol=1e-10)
    assert np.allclose(out[capi.RIEMANN_DISSIPATIVE][1], V * rho0 * c0 * u * dW * V * e01, rtol=1e-10)


@pytest.mark.parametrize("dim", [2, 3])
def test_observer_interpolation_closed_form(oracle, dim):
    """Interpolation<Contact<Real>>: p_obs = sum W V p / (sum W V + TinyReal)
    (interpolation_dynamics.hpp:46-60), with the targets at table nodes so
    that the tabulated W equals the closed form."""
    dp = 0.1
    h = 1.3 * dp
    nodes = [3, 7, 12, 16]
    tpos = np.zeros((len(nodes), dim))
    for k, q in enumerate(nodes):
        tpos[k, k % dim] = (1 if k < 2 else -1) * q * 0.1 * h
    ctx = ck.Context(oracle)
    tb = ck.Body(ctx, dim, len(nodes), dp)
    tb.write("Position", tpos)
    vol = np.array([1.0, 0.8, 1.2, 0.9]) * dp ** dim
    tb.write("VolumetricMeasure", vol)
    C = np.array([1.001, 1.002, 0.999, 1.0005])
    tb.write("Compression", C)
    ob = ck.Body(ctx, dim, 1, dp, capi.BODY_OBSERVER)
    ob.write("Position", np.zeros((1, dim)))
    lo, up = domain(np.vstack([tpos, np.zeros((1, dim))]))
    tc = ck.CellLinkedList(tb, lo, up)
    rel = ck.Contact(ob, tb, tc)
    ck.UpdateCellLinkedList(tc).exec()
    ck.UpdateRelation(rel).exec()
    p = 1.0 * 1.0 ** 2 * (C - 1.0)  # rho0 c0^2 (C - 1)
    tb.write("Pressure", p)
    w = dimfac(dim) / h ** dim * np.array([wendland_w(q * 0.1) for q in nodes]) * vol
    got = ck.InterpolationContact(rel, "Pressure").exec()[0]
    assert got == pytest.approx(np.sum(w * p) / (np.sum(w) + np.finfo(np.float64).eps), rel=1e-12)
    ctx.close()


def test_fp32_restatement_against_fp64(oracle):
    """The FP32 restatement (the reference's SPHINXSYS_USE_FLOAT build) on the
    same 3D scene as the pinned FP64 one: neighbour lists identical away from
    the cut-off, summation and pair forces within float rounding.  The
    compression rate sums p_i - p_j of neighbours whose C differs by ~1e-5, so
    float storage of C (6e-8 absolute) costs it ~1e-4 relative: gate 1e-3."""
    rng = np.random.default_rng(12)
    dp = 0.05
    m = 9
    axes = [np.arange(m) * dp for _ in range(3)]
    pos = np.stack([g.ravel() for g in np.meshgrid(*axes, indexing="ij")], axis=1)
    pos += rng.uniform(-0.1 * dp, 0.1 * dp, size=pos.shape)
    res = {}
    for prec in (capi.FP64, 2):
        ctx, b = make_body(oracle, pos, dp, rho0=1000.0, c0=20.0, precision=prec)
        lo, up = domain(pos)
        cll = ck.CellLinkedList(b, lo, up)
        inner = ck.Inner(b, cll)
        ck.UpdateCellLinkedList(cll).exec()
        ck.UpdateRelation(inner).exec()
        b.write("Compression", 1.0 + np.linspace(-0.005, 0.005, len(pos)))
        b.write("Velocity", np.outer(np.linspace(-1, 1, len(pos)), np.ones(3)) * 0.1)
        b.write("Mass", np.full(len(pos), 1000.0 * dp ** 3))
        ck.CompressionSummation(inner).exec()
        ck.AcousticStep1stHalf(inner).exec(1e-5)
        res[prec] = (inner.read(), b.read("CompressionSummation"), b.read("Force"), b.read("CompressionRate"))
        ctx.close()
    (r64, s64, f64, c64), (r32, s32, f32, c32) = res[capi.FP64], res[2]
    for a, b_ in zip(r64, r32):
        assert np.array_equal(a, b_)
    rel = lambda a, b_: np.linalg.norm(a - b_) / np.linalg.norm(b_)
    assert rel(s32, s64) < 1e-6
    assert rel(f32, f64) < 1e-4
    assert rel(c32, c64) < 1e-3
