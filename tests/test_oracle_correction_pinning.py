"""Pin the oracle's linear kernel correction and free-surface indication.

Checks are independent of the oracle's code: the correction matrix against
numpy linear algebra (np.linalg.det / inv) on neighbours placed at table
nodes (where the tabulated kernel equals the Wendland C2 closed form) and
against the lattice closed form; the corrected half steps against the
hand-derived two-particle formulas of acoustic_step_{1st,2nd}_half.hpp with
non-symmetric matrices (so that a B_i/B_j swap or a transpose fails); the
surface indication against a brute-force restatement of its rule with the
closed-form kernel on scenes whose decisions have a margin."""
import math

import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import ck
from tests.test_oracle_pinning import dimfac, domain, make_body, wendland_dw

SQRT_EPS = 1.0e-8


def grad_w(q, dim, h):
    return dimfac(dim) / h ** (dim + 1) * wendland_dw(q)


def np_update(Braw, alpha):
    """kernel_correction_ck.hpp:58-65 in numpy linear algebra."""
    det = np.linalg.det(Braw)
    det_sqr = max(alpha - det, 0.0)
    inv = np.linalg.inv(Braw.T @ Braw + SQRT_EPS * np.eye(len(Braw))) @ Braw.T
    w = det / (det + det_sqr)
    return w * inv + (1.0 - w) * np.eye(len(Braw))


def scene(lib, fluid, dp=0.1, wall=None, rho0=1.0, c0=10.0):
    ctx = ck.Context(lib)
    dim = fluid.shape[1]
    fb = ck.Body(ctx, dim, len(fluid), dp, rho0=rho0, c0=c0)
    fb.write("Position", fluid)
    fb.write("VolumetricMeasure", np.full(len(fluid), dp ** dim))
    rels = []
    allpos = fluid if wall is None else np.vstack([fluid, wall])
    lo, up = domain(allpos)
    fc = ck.CellLinkedList(fb, lo, up)
    inner = ck.Inner(fb, fc)
    ck.UpdateCellLinkedList(fc).exec()
    wb = None
    if wall is not None:
        wb = ck.Body(ctx, dim, len(wall), dp, capi.BODY_SOLID)
        wb.write("Position", wall)
        wb.write("VolumetricMeasure", np.full(len(wall), dp ** dim))
        wc = ck.CellLinkedList(wb, lo, up)
        ck.UpdateCellLinkedList(wc).exec()
        rels.append(ck.Contact(fb, wb, wc))
    ck.UpdateRelation(inner, *rels).exec()
    return ctx, fb, wb, inner, rels


def node_star(dim, rng, h, qs):
    """Particle 0 at the origin, one neighbour per q in qs at distance q h in a
    random direction (all pair distances of particle 0 on table nodes)."""
    pts = [np.zeros(dim)]
    for q in qs:
        d = rng.normal(size=dim)
        pts.append(q * h * d / np.linalg.norm(d))
    return np.array(pts)


# ---------------------------------------------------------------- correction matrix
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("alpha", [0.0, 0.5, 5.0])
def test_correction_matrix_at_table_nodes_against_numpy(oracle, dim, alpha):
    """B_0 = update(-sum r_0j (x) dW(r_0j) e_0j V_j) for neighbours on table
    nodes, with alpha = 0 (pure Tikhonov inverse), 0.5 and 5 (blended)."""
    dp = 0.1
    h = 1.3 * dp
    rng = np.random.default_rng(11 + dim)
    qs = [0.3, 0.5, 0.8, 1.1, 1.4, 1.7, 0.6, 1.2] if dim == 3 else [0.4, 0.7, 1.0, 1.5, 0.9]
    pos = node_star(dim, rng, h, qs)
    ctx, fb, _, inner, _ = scene(oracle, pos, dp)
    ck.LinearCorrectionMatrixComplex(inner, alpha=alpha).exec()
    B = fb.read("LinearCorrectionMatrix")
    V = dp ** dim
    Braw = np.zeros((dim, dim))
    for j, q in enumerate(qs, start=1):
        r = pos[0] - pos[j]
        e = r / np.linalg.norm(r)
        Braw -= np.outer(r, grad_w(q, dim, h) * e * V)
    expect = np_update(Braw, alpha)
    w = np.linalg.det(Braw) / (np.linalg.det(Braw) + max(alpha - np.linalg.det(Braw), 0.0))
    if alpha == 5.0:
        assert 0.0 < w < 1.0  # the blended branch is exercised
    assert np.allclose(B[0], expect, rtol=1e-9, atol=1e-12)
    ctx.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_correction_matrix_interior_lattice_closed_form(oracle, dim):
    """On a perfect lattice B_raw = c I with c = -(1/dim) sum |r| dW V (closed-form
    Wendland sum, tabulation error only); B = c/(c^2 + 1e-8) I."""
    dp = 0.1
    n = 9 if dim == 3 else 13
    axes = [np.arange(n) * dp for _ in range(dim)]
    pos = np.stack([g.ravel() for g in np.meshgrid(*axes, indexing="ij")], axis=1)
    ctx, fb, _, inner, _ = scene(oracle, pos, dp)
    ck.LinearCorrectionMatrixComplex(inner, alpha=0.5).exec()
    B = fb.read("LinearCorrectionMatrix")
    centre = np.argmin(np.sum((pos - pos.mean(axis=0)) ** 2, axis=1))
    h = 1.3
    s = 0.0
    for idx in np.ndindex(*([7] * dim)):
        rr = math.sqrt(sum((k - 3) ** 2 for k in idx))
        q = rr / h
        if 0 < q < 2.0:
            s -= rr * wendland_dw(q)
    c = s * dimfac(dim) / h ** (dim + 1) / dim  # in units of dp: |r| dW V scales as dp^0
    assert abs(c - 1.0) < 0.05  # the kernel gradient is nearly consistent on a lattice
    expect = c / (c * c + SQRT_EPS) * np.eye(dim)
    assert np.allclose(B[centre], expect, rtol=2e-4, atol=2e-5)
    ctx.close()


def test_correction_matrix_isolated_particle_is_identity(oracle):
    """No neighbour: B_raw = 0, det = 0, weight = 0 (alpha > 0), B = I."""
    pos = np.array([[0.0, 0.0, 0.0], [5.0, 5.0, 5.0]])
    ctx, fb, _, inner, _ = scene(oracle, pos, 0.1)
    fb.write("LinearCorrectionMatrix", np.full((2, 3, 3), 7.0))
    ck.LinearCorrectionMatrixComplex(inner, alpha=0.5).exec()
    assert np.array_equal(fb.read("LinearCorrectionMatrix"), np.tile(np.eye(3), (2, 1, 1)))
    ctx.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_correction_matrix_wall_contact_part(oracle, dim):
    """The contact interaction adds -sum r_iw (x) dW e V_w for the wall neighbours."""
    dp = 0.1
    h = 1.3 * dp
    rng = np.random.default_rng(3)
    fluid = node_star(dim, rng, h, [0.5, 1.2])
    wall = node_star(dim, rng, h, [0.9, 1.6])[1:]
    ctx, fb, _, inner, rels = scene(oracle, fluid, dp, wall)
    ck.LinearCorrectionMatrixComplex(inner, *rels, alpha=0.0).exec()
    V = dp ** dim
    Braw = np.zeros((dim, dim))
    for tar, qs in ((fluid[1:], [0.5, 1.2]), (wall, [0.9, 1.6])):
        for x, q in zip(tar, qs):
            r = fluid[0] - x
            Braw -= np.outer(r, grad_w(q, dim, h) * r / np.linalg.norm(r) * V)
    assert np.allclose(fb.read("LinearCorrectionMatrix")[0], np_update(Braw, 0.0), rtol=1e-9, atol=1e-12)
    ctx.close()


# ---------------------------------------------------------------- corrected half steps
def pair_scene(lib, dim, q, rho0=1.0, c0=10.0):
    dp = 0.1
    h = 1.3 * dp
    pos = np.zeros((2, dim))
    pos[1] = q * h * np.ones(dim) / math.sqrt(dim)  # diagonal: every component of e is non-zero
    ctx, fb, _, inner, _ = scene(lib, pos, dp, rho0=rho0, c0=c0)
    e01 = (pos[0] - pos[1]) / np.linalg.norm(pos[0] - pos[1])
    return ctx, fb, inner, grad_w(q, dim, h), dp ** dim, e01


def random_mats(rng, n, dim):
    return np.eye(dim) + 0.3 * rng.normal(size=(n, dim, dim))  # not symmetric


@pytest.mark.parametrize("dim", [2, 3])
def test_corrected_first_half_two_particles(oracle, dim):
    """F_i = -V_i (p_i B_j + p_j B_i) e_ij dW V_j (AverageP of one fluid is the mean);
    the continuity term has no correction; F_0 + F_1 = 0."""
    rho0, c0 = 1.0, 10.0
    ctx, fb, inner, dW, V, e01 = pair_scene(oracle, dim, 1.1, rho0, c0)
    rng = np.random.default_rng(7)
    B = random_mats(rng, 2, dim)
    fb.write("LinearCorrectionMatrix", B)
    C0 = np.array([1.004, 1.001])
    fb.write("Compression", C0)
    fb.write("Mass", np.full(2, 1.0))
    ck.AcousticStep1stHalfWithWallRiemannCorrectionCK(inner).exec(0.0)
    p = rho0 * c0 ** 2 * (C0 - 1.0)
    F = fb.read("Force")
    expect0 = -V * (p[0] * B[1] + p[1] * B[0]) @ e01 * dW * V
    expect1 = -V * (p[1] * B[0] + p[0] * B[1]) @ (-e01) * dW * V
    assert np.allclose(F[0], expect0, rtol=1e-10, atol=0)
    assert np.allclose(F[1], expect1, rtol=1e-10, atol=0)
    assert np.allclose(F[0] + F[1], 0.0, atol=1e-12 * np.abs(F).max())
    dC = fb.read("CompressionRate")
    assert dC[0] == pytest.approx(C0[0] * (p[0] - p[1]) / (rho0 * c0) * dW * V, rel=1e-10)
    ctx.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_corrected_second_half_two_particles(oracle, dim):
    """Ċ_i = C_i (v_i - v_j).(B_i e_ij) dW V_j; the dissipation force keeps e_ij."""
    rho0, c0 = 1.0, 10.0
    ctx, fb, inner, dW, V, e01 = pair_scene(oracle, dim, 0.7, rho0, c0)
    rng = np.random.default_rng(8)
    B = random_mats(rng, 2, dim)
    fb.write("LinearCorrectionMatrix", B)
    vel = rng.normal(size=(2, dim))
    fb.write("Velocity", vel)
    ck.AcousticStep2ndHalfWithWallRiemannCorrectionCK(inner).exec(0.0)
    dC = fb.read("CompressionRate")
    assert dC[0] == pytest.approx(np.dot(vel[0] - vel[1], B[0] @ e01) * dW * V, rel=1e-10)
    assert dC[1] == pytest.approx(np.dot(vel[1] - vel[0], B[1] @ -e01) * dW * V, rel=1e-10)
    u = np.dot(vel[0] - vel[1], e01)
    pj = rho0 * c0 * u * min(3.0 * max(u, 0.0) / c0, 1.0)
    assert np.allclose(fb.read("Force")[0], V * pj * dW * V * e01, rtol=1e-10, atol=1e-300)
    ctx.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_corrected_wall_terms(oracle, dim):
    """Wall: F_i -= (p_i + p_w) B_i e dW V_w V_i; Ċ_i += C_i 2(v_i - v_w).(B_i e) dW V_w."""
    rho0, c0 = 1.0, 10.0
    dp = 0.1
    h = 1.3 * dp
    q = 0.9
    d = np.ones(dim) / math.sqrt(dim)
    fluid = np.array([q * h * d])
    wall = np.zeros((1, dim))
    ctx, fb, wb, inner, rels = scene(oracle, fluid, dp, wall, rho0, c0)
    n = np.zeros((1, dim))
    n[0, -1] = 1.0
    wb.write("NormalDirection", n)
    rng = np.random.default_rng(9)
    B = random_mats(rng, 1, dim)
    fb.write("LinearCorrectionMatrix", B)
    m = 0.7
    fb.write("Mass", np.full(1, m))
    g = np.zeros((1, dim))
    g[0, -1] = -m
    fb.write("ForcePrior", g)
    C0 = 1.003
    fb.write("Compression", np.full(1, C0))
    ck.AcousticStep1stHalf(inner, *rels, correction=capi.CORRECTION_LINEAR).exec(0.0)
    V = dp ** dim
    dW = grad_w(q, dim, h)
    p_i = rho0 * c0 ** 2 * (C0 - 1.0)
    e = d
    p_w = p_i + C0 * rho0 * q * h * max(0.0, np.dot(g[0] / m, -e))
    assert np.allclose(fb.read("Force")[0], -(p_i + p_w) * (B[0] @ e) * dW * V * V, rtol=1e-10)
    vel = rng.normal(size=(1, dim))
    fb.write("Velocity", vel)
    fb.write("CompressionRate", np.zeros(1))
    ck.AcousticStep2ndHalf(inner, *rels, correction=capi.CORRECTION_LINEAR).exec(0.0)
    assert fb.read("CompressionRate")[0] == pytest.approx(
        C0 * np.dot(2.0 * vel[0], B[0] @ e) * dW * V, rel=1e-10)
    ctx.close()


def jittered_block(dim, m, dp, seed, amp=0.15):
    rng = np.random.default_rng(seed)
    axes = [np.arange(m) * dp for _ in range(dim)]
    pos = np.stack([g.ravel() for g in np.meshgrid(*axes, indexing="ij")], axis=1)
    return pos + rng.uniform(-amp * dp, amp * dp, size=pos.shape), rng


@pytest.mark.parametrize("dim", [2, 3])
def test_identity_correction_equals_uncorrected(oracle, dim):
    """LinearCorrectionCK with B = I gives the NoKernelCorrectionCK result."""
    dp = 0.05
    pos, rng = jittered_block(dim, 10 if dim == 2 else 6, dp, 21)
    out = []
    for corr in (capi.CORRECTION_NONE, capi.CORRECTION_LINEAR):
        ctx, fb, _, inner, _ = scene(oracle, pos, dp)
        fb.write("Compression", 1.0 + np.random.default_rng(1).uniform(0, 0.02, len(pos)))
        fb.write("Velocity", np.random.default_rng(2).normal(size=pos.shape))
        fb.write("Mass", np.full(len(pos), 1.0))
        ck.AcousticStep1stHalf(inner, correction=corr).exec(1e-4)
        F1 = fb.read("Force")
        ck.AcousticStep2ndHalf(inner, correction=corr).exec(1e-4)
        out.append((F1, fb.read("Force"), fb.read("Compression")))
        ctx.close()
    for a, b in zip(*out):
        assert np.allclose(a, b, rtol=1e-14, atol=1e-14 * np.abs(a).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_corrected_pressure_force_conserves_momentum(oracle, dim):
    """sum_i F_i = 0: the pair term -(p_i B_j + p_j B_i) e_ij is antisymmetric in (i, j)."""
    dp = 0.05
    pos, rng = jittered_block(dim, 12 if dim == 2 else 7, dp, 5)
    ctx, fb, _, inner, _ = scene(oracle, pos, dp)
    ck.LinearCorrectionMatrixComplex(inner, alpha=0.5).exec()
    B = fb.read("LinearCorrectionMatrix")
    assert np.abs(B - np.eye(dim)).max() > 0.05  # the surface rows are corrected
    fb.write("Compression", 1.0 + rng.uniform(0, 0.02, len(pos)))
    fb.write("Mass", np.full(len(pos), 1.0))
    ck.AcousticStep1stHalfWithWallRiemannCorrectionCK(inner).exec(0.0)
    F = fb.read("Force")
    assert np.abs(F.sum(axis=0)).max() < 1e-12 * np.abs(F).sum()
    ctx.close()


# ---------------------------------------------------------------- free-surface indication
def np_surface_indication(fluid, wall, dp, prev):
    """surface_indication_ck.hpp:55-127 restated with brute-force neighbours and
    the closed-form kernel gradient: returns (pos_div, indicator, margin), where
    margin is the smallest distance of a compared value to the threshold."""
    dim = fluid.shape[1]
    h = 1.3 * dp
    V = dp ** dim
    thr = 0.75 * dim

    def sums(src, tar, inner):
        out, nbrs = [], []
        for i in range(len(src)):
            r = np.linalg.norm(src[i] - tar, axis=1)
            m = r < 2.0 * h
            if inner:
                m[i] = False
            idx = np.nonzero(m)[0]
            nbrs.append((idx, r[idx]))
            out.append(-sum(grad_w(rr / h, dim, h) * V * rr for rr in r[idx]))
        return np.array(out), nbrs

    pd, nb = sums(fluid, fluid, True)
    margins = list(np.abs(pd - thr))
    for i in range(len(fluid)):
        if pd[i] < thr and prev[i] != 1 and not np.any(prev[nb[i][0]] == 1):
            pd[i] = 2.0 * thr
    if wall is not None:
        pd = pd + sums(fluid, wall, False)[0]
    margins += list(np.abs(pd - thr))
    ind = np.ones(len(fluid), dtype=np.int32)
    for i in range(len(fluid)):
        idx, r = nb[i]
        margins += list(np.abs(r - h))
        if pd[i] > thr and not np.any((pd[idx] < thr) & (r < h)):
            ind[i] = 0
    return pd, ind, min(margins)


@pytest.mark.parametrize("dim", [2, 3])
def test_surface_indication_lattice_block_spatial(oracle, dim):
    """First call (PreviousSurfaceIndicator = 1 everywhere): spatial rule only;
    interior divergence near dim, the outer layers flagged."""
    dp = 0.1
    n = 8 if dim == 2 else 6
    axes = [np.arange(n) * dp for _ in range(dim)]
    pos = np.stack([g.ravel() for g in np.meshgrid(*axes, indexing="ij")], axis=1)
    ctx, fb, _, inner, _ = scene(oracle, pos, dp)
    ck.FreeSurfaceIndicationComplexSpatialTemporalCK(inner).exec()
    pd_np, ind_np, margin = np_surface_indication(pos, None, dp, np.ones(len(pos), np.int32))
    assert margin > 1e-3
    pd = fb.read("PositionDivergence")
    assert np.allclose(pd, pd_np, rtol=2e-4, atol=1e-6)
    ind = fb.read("Indicator")
    assert np.array_equal(ind, ind_np)
    assert np.array_equal(fb.read("PreviousSurfaceIndicator"), ind)
    centre = np.argmin(np.sum((pos - pos.mean(axis=0)) ** 2, axis=1))
    assert abs(pd[centre] - dim) < 0.05 * dim and ind[centre] == 0
    assert ind[0] == 1  # a corner
    ctx.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_surface_indication_temporal_rule_and_contact_order(oracle, dim):
    """With PreviousSurfaceIndicator 0 except near one corner, a low inner
    divergence far from that corner is replaced by 2 x threshold before the wall
    part is added; indicator and divergence against the brute-force rule."""
    dp = 0.1
    n = 8 if dim == 2 else 5
    rng = np.random.default_rng(7)  # a jitter whose decisions all have a margin above 2e-3
    axes = [np.arange(n) * dp for _ in range(dim)]
    pos = np.stack([g.ravel() for g in np.meshgrid(*axes, indexing="ij")], axis=1)
    pos = pos + rng.uniform(-0.05 * dp, 0.05 * dp, size=pos.shape)
    wall_axes = [np.arange(-1, n + 1) * dp for _ in range(dim - 1)]
    wall2 = np.stack([g.ravel() for g in np.meshgrid(*wall_axes, indexing="ij")], axis=1)
    wall = np.hstack([wall2, np.full((len(wall2), 1), -dp)])  # one layer below the block
    ctx, fb, _, inner, rels = scene(oracle, pos, dp, wall)
    prev = np.zeros(len(pos), np.int32)
    prev[0] = 1
    fb.write("PreviousSurfaceIndicator", prev)
    ck.FreeSurfaceIndicationComplexSpatialTemporalCK(inner, *rels).exec()
    pd_np, ind_np, margin = np_surface_indication(pos, wall, dp, prev)
    assert margin > 1e-3
    pd = fb.read("PositionDivergence")
    assert np.allclose(pd, pd_np, rtol=2e-4, atol=1e-6)
    assert np.array_equal(fb.read("Indicator"), ind_np)
    thr = 0.75 * dim
    forced = pd > 2.0 * thr  # held at 2 thr, then the wall part added on top
    assert forced.any()
    assert ind_np.sum() < len(pos)
    ctx.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_previous_indicator_is_evolving_under_sort(oracle, dim):
    """PreviousSurfaceIndicator follows the particles through ParticleSortCK;
    Indicator (not evolving) stays in place."""
    dp = 0.1
    pos, rng = jittered_block(dim, 6 if dim == 2 else 4, dp, 8, 0.3)
    pos = pos[rng.permutation(len(pos))]
    ctx, fb, _, inner, _ = scene(oracle, pos, dp)
    prev = rng.integers(0, 2, len(pos)).astype(np.int32)
    ind = rng.integers(0, 2, len(pos)).astype(np.int32)
    fb.write("PreviousSurfaceIndicator", prev)
    fb.write("Indicator", ind)
    ck.ParticleSortCK(fb, inner.cll).exec()
    oid = fb.read("OriginalID")
    assert not np.array_equal(oid, np.arange(len(pos)))
    assert np.array_equal(fb.read("PreviousSurfaceIndicator"), prev[oid])
    assert np.array_equal(fb.read("Indicator"), ind)
    ctx.close()
