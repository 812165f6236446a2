"""Independent scenes served together on one GPU: each result is the same
alone and in the mix.  Every scene owns its context, stream and device state
and every kernel is deterministic, so the fluid state after the run is
compared bit for bit."""
import numpy as np
import pytest

from sphinxsys_amd import serving

pytestmark = pytest.mark.gpu


def test_scenes_alone_and_in_the_mix(hip):
    specs = serving.batched_2d_specs(count=6, n_lo=1000, n_hi=20000, steps=25, seed=7)
    alone, _ = serving.Server(hip, max_active=1).run(specs, keep_fields=True, realtime=False)
    mixed, _ = serving.Server(hip, max_active=6).run(specs, keep_fields=True, realtime=False)
    for a, b in zip(alone, mixed):
        assert a.acoustic_steps == b.acoustic_steps and a.physical_time == b.physical_time
        for k in a.fields:
            assert np.array_equal(a.fields[k], b.fields[k]), (a.ident, k)


def test_admission_respects_slots_and_budget(hip):
    specs = serving.batched_2d_specs(count=5, n_lo=1000, n_hi=4000, steps=5, seed=3)
    budget = max(s.footprint_bytes() for s in specs) * 2
    res, _ = serving.Server(hip, max_active=4, hbm_budget=budget).run(specs, realtime=False)
    # at no time more than two scenes were resident: intervals overlap at most pairwise
    spans = sorted((r.start, r.finish) for r in res)
    for t in [s for s, _ in spans]:
        assert sum(1 for s, f in spans if s <= t < f) <= 2
    assert all(r.acoustic_steps > 0 for r in res)


def test_served_3d_stream_alone_and_in_the_batch(hip):
    """A Poisson stream of 3D dam breaks served by continuous batching on the
    batch path (scenes join and leave the batch as they arrive and finish,
    each sorting every 100 advection steps as its own loop does): every
    scene's fields, step counts and physical time are bit-identical to the
    same request run alone."""
    specs = serving.served_3d_specs(count=5, rate=40.0, n_lo=20000, n_hi=150000, t_end=0.15, seed=9)
    alone, _ = serving.Server(hip, max_active=1).run(specs, keep_fields=True, realtime=False)
    server = serving.BatchServer(hip, max_active=3)
    mixed, _ = server.run(specs, keep_fields=True, realtime=True)
    assert max(r.advection_steps for r in alone) >= 100  # a sort inside the runs
    for a, b in zip(alone, mixed):
        assert (a.acoustic_steps, a.advection_steps, a.physical_time) == \
            (b.acoustic_steps, b.advection_steps, b.physical_time), a.ident
        for k in a.fields:
            assert np.array_equal(a.fields[k], b.fields[k]), (a.ident, k)
    # with three slots for five requests, scenes were admitted as others left
    assert server.tuning["steps"] > max(r.advection_steps for r in mixed)


def test_served_3d_stream_20k_to_2m_to_t_end_1(hip):
    """The served configuration at full size: four 3D dam breaks of 28k to 1.8M
    fluid particles (log-uniform in 20k-2M) arriving at 4 requests/s, each run
    to t = 1.0 s, served by continuous batching with two slots (later requests
    wait for a slot to free); every scene's fields, step counts and physical
    time are bit-identical to its run alone."""
    specs = serving.served_3d_specs(count=4, rate=4.0, n_lo=20000, n_hi=2000000, t_end=1.0, seed=13)
    alone, _ = serving.Server(hip, max_active=1).run(specs, keep_fields=True, realtime=False)
    assert max(r.n_fluid for r in alone) > 1500000 and min(r.n_fluid for r in alone) < 50000
    assert all(r.physical_time >= 1.0 for r in alone)
    server = serving.BatchServer(hip, max_active=2)
    mixed, _ = server.run(specs, keep_fields=True, realtime=True)
    for a, b in zip(alone, mixed):
        assert (a.acoustic_steps, a.advection_steps, a.physical_time) == \
            (b.acoustic_steps, b.advection_steps, b.physical_time), a.ident
        for k in a.fields:
            assert np.array_equal(a.fields[k], b.fields[k]), (a.ident, k)


def test_batched_2d_scenes_alone_and_in_the_batch(hip):
    specs = serving.batched_2d_specs(count=12, n_lo=1000, n_hi=20000, steps=25, seed=7)
    alone, _ = serving.Server(hip, max_active=1).run(specs, keep_fields=True, realtime=False)
    mixed, _ = serving.BatchServer(hip, max_active=5).run(specs, keep_fields=True, realtime=False)
    for a, b in zip(alone, mixed):
        assert a.acoustic_steps == b.acoustic_steps and a.physical_time == b.physical_time
        for k in a.fields:
            assert np.array_equal(a.fields[k], b.fields[k]), (a.ident, k)
