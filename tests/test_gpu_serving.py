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
