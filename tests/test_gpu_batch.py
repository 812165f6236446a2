"""Batched scenes (sphx_batch_*) on the GPU.

Each scene advanced by the batch -- one launch per stage over all scenes,
the step captured in a HIP graph and replayed -- must end bit-identical to
the same scene advanced alone with sphx_advection_step: positions,
velocities, compression, forces, particle order, physical time and sub-step
counts.  Covered: FP64 and FP32-mixed, graph and direct launches, the
autotuner's shape changes (LDS staging switched per scene, tiled-build waves
per workgroup), fewer captured sub-steps than a scene needs (the host
finishes it), a particle sort inside the batch, a diverged scene, and the
256-tank batched workload (1k-64k particles per tank) over 40 advection
steps.  FP64 batched scenes are also checked against the oracle."""
import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import cases, ck, serving

pytestmark = pytest.mark.gpu

FIELDS = ("Position", "Velocity", "Compression", "Force", "Density", "OriginalID")


def make(lib, spec, precision):
    sc = cases.DamBreak(lib, spec.geometry(), precision)
    sc.initialize()
    return sc


def snapshot(sc):
    return {k: sc.fluid.read(k) for k in FIELDS}


def run_alone(lib, specs, precision, steps, sort_at=None, poison=None):
    """Every scene alone through sphx_advection_step; per scene (fields, time, sub-step counts)."""
    out = []
    for k, spec in enumerate(specs):
        sc = make(lib, spec, precision)
        if poison is not None and k == poison:
            v = sc.fluid.read("Velocity")
            v[len(v) // 2, 0] = np.nan
            sc.fluid.write("Velocity", v)
        t, counts = 0.0, []
        try:
            for s in range(steps):
                sort = sort_at is not None and sort_at(k, s)
                t, n = sc.advection.exec(t, sort_now=sort)
                counts.append(n)
            out.append((snapshot(sc), t, counts))
        except capi.SphxError as e:
            assert "status 4" in str(e)  # SPHX_E_NONFINITE
            out.append(None)
        finally:
            sc.close()
    return out


def run_batched(lib, specs, precision, steps, options=None, sort_at=None, poison=None):
    scenes = [make(lib, spec, precision) for spec in specs]
    if poison is not None:
        v = scenes[poison].fluid.read("Velocity")
        v[len(v) // 2, 0] = np.nan
        scenes[poison].fluid.write("Velocity", v)
    batch = ck.BatchAdvection(lib, [sc.advection for sc in scenes], options=options)
    counts = [[] for _ in specs]
    statuses = []
    try:
        for s in range(steps):
            sort = None
            if sort_at is not None:
                sort = [1 if sort_at(k, s) else 0 for k in range(len(specs))]
            st = batch.exec(sort_now=sort)
            statuses.append(st)
            for k in range(len(specs)):
                counts[k].append(int(batch.steps[k]))
        out = []
        for k, sc in enumerate(scenes):
            out.append((snapshot(sc), float(batch.time[k]), counts[k]))
        return out, np.array(statuses), batch.tuning()
    finally:
        batch.close()
        for sc in scenes:
            sc.close()


def assert_same(alone, batched, skip=()):
    for k, (a, b) in enumerate(zip(alone, batched)):
        if k in skip:
            continue
        fa, ta, ca = a
        fb, tb, cb = b
        assert ca == cb, (k, ca, cb)
        assert ta == tb, (k, ta, tb)
        for name in FIELDS:
            assert np.array_equal(fa[name], fb[name]), (k, name)


def specs_small(count=10, seed=11, steps=8):
    return serving.batched_2d_specs(count, seed=seed, n_lo=400, n_hi=9000, steps=steps)


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64])
def test_batch_graph_autotuned_bit_identical_to_alone(hip, precision):
    specs = specs_small()
    steps = 9  # first step direct, then the tuner's six shapes, then the chosen one
    print(f"batch of {len(specs)} 2D scenes, precision {precision}, autotuned graph", flush=True)
    alone = run_alone(hip, specs, precision, steps)
    batched, st, tuning = run_batched(hip, specs, precision, steps)
    assert (st == 0).all()
    assert_same(alone, batched)
    assert tuning["graph_launches"] == steps - 1 and tuning["direct_steps"] == 1
    if precision == capi.MIXED:
        assert tuning["tuned"] and len(tuning["candidates"]) == 6
        assert all(c["ms_per_million_particle_substeps"] is not None for c in tuning["candidates"])


@pytest.mark.parametrize("use_graph,margin", [(0, 1), (1, 0)])
def test_batch_direct_and_short_graph(hip, use_graph, margin):
    """Direct launches; and a graph with no spare sub-step, so that scenes
    needing one more sub-step than the last step are finished by the host."""
    specs = specs_small(count=8, seed=12, steps=6)
    steps = 6
    opts = dict(use_graph=use_graph, autotune=0, tile_waves=2, stage_min=3000, substep_margin=margin)
    print(f"batch use_graph={use_graph} margin={margin}", flush=True)
    alone = run_alone(hip, specs, capi.MIXED, steps)
    batched, st, tuning = run_batched(hip, specs, capi.MIXED, steps, options=opts)
    assert (st == 0).all()
    assert_same(alone, batched)
    if use_graph:
        assert tuning["graph_launches"] == steps - 1
    else:
        assert tuning["direct_steps"] == steps


def test_batch_particle_sort_in_the_mix(hip):
    specs = specs_small(count=6, seed=13, steps=5)
    steps = 5

    def sort_at(k, s):
        return s == 2 and k % 2 == 0

    print("batch with sorts at step 2", flush=True)
    alone = run_alone(hip, specs, capi.MIXED, steps, sort_at=sort_at)
    batched, st, tuning = run_batched(hip, specs, capi.MIXED, steps, sort_at=sort_at)
    assert (st == 0).all()
    assert_same(alone, batched)
    assert tuning["host_tail_scenes"] >= 3


def test_batch_diverged_scene_reported_others_unaffected(hip):
    specs = specs_small(count=5, seed=14, steps=3)
    steps = 3
    print("batch with a poisoned scene", flush=True)
    alone = run_alone(hip, specs, capi.MIXED, steps, poison=2)
    assert alone[2] is None
    batched, st, _ = run_batched(hip, specs, capi.MIXED, steps, poison=2)
    assert (st[:, 2] == 4).all()  # SPHX_E_NONFINITE every step
    assert (np.delete(st, 2, axis=1) == 0).all()
    assert_same(alone, batched, skip=(2,))


def test_batch_fp64_against_oracle(hip, oracle):
    specs = serving.batched_2d_specs(3, seed=15, n_lo=500, n_hi=2500, steps=4)
    steps = 4
    print("FP64 batch against the oracle", flush=True)
    batched, st, _ = run_batched(hip, specs, capi.FP64, steps)
    assert (st == 0).all()
    ref, st_ref, _ = run_batched(oracle, specs, capi.FP64, steps)
    assert (st_ref == 0).all()
    for (fb, tb, cb), (fr, tr, cr) in zip(batched, ref):
        assert cb == cr
        assert abs(tb - tr) <= 1e-12 * max(abs(tr), 1.0)
        assert np.array_equal(fb["OriginalID"], fr["OriginalID"])
        for name in ("Position", "Velocity", "Compression"):
            a, b = fb[name], fr[name]
            err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
            assert err < 1e-9, (name, err)


def test_batched_256_tanks_40_steps_bit_identical(hip):
    """The batched workload: 256 independent 2D tanks (1k-64k fluid particles,
    log-uniform), 40 advection steps, FP32-mixed, one graph launch per step."""
    specs = serving.batched_2d_specs(256, steps=40)
    steps = 40
    print(f"256 tanks, {sum(s.fluid_count() for s in specs)} fluid particles", flush=True)
    batched, st, tuning = run_batched(hip, specs, capi.MIXED, steps)
    assert (st == 0).all()
    assert tuning["graph_launches"] == steps - 1
    print(f"batched done: {tuning}", flush=True)
    alone = run_alone(hip, specs, capi.MIXED, steps)
    assert_same(alone, batched)


def test_batch_3d_unstaged_bit_identical_to_staged_alone(hip):
    """A 3D FP32-mixed dam break (dp 0.02, 90k particles) whose blocks range
    above and below the passes' staging caps: in a batch with LDS staging
    switched off (every neighbour read from global memory) it stays
    bit-identical to the scene alone (staged) over 150 advection steps with a
    particle sort.  The pair arithmetic of the staged and global neighbour
    walks is written out (explicit multiply-adds), so no walk depends on how
    the compiler fuses a sum of products."""
    geom = cases.dambreak_3d(dp=0.02)
    alone = cases.DamBreak(hip, geom, capi.MIXED)
    alone.initialize()
    other = cases.DamBreak(hip, geom, capi.MIXED)
    other.initialize()
    batch = ck.BatchAdvection(hip, [other.advection], options=dict(use_graph=1, autotune=0, stage_min=1 << 30))
    try:
        for s in range(150):
            alone.advection_step()
            batch.exec(sort_now=[1 if (s + 1) % 100 == 0 else 0])
        assert alone.iterations == 150 and batch.time[0] == alone.time
        for k in FIELDS:
            assert np.array_equal(alone.fluid.read(k), other.fluid.read(k)), k
    finally:
        batch.close()
        alone.close()
        other.close()
