"""Host logic of the scene server, driven on the oracle with tiny scenes:
the request stream, admission by slots and HBM budget, and that a scene's
result is the same alone and in the mix."""
import numpy as np

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import serving


def test_specs_follow_the_stated_distributions():
    specs = serving.batched_2d_specs(count=256)
    n = np.array([s.fluid_count() for s in specs])
    assert n.min() >= 900 and n.max() <= 66000
    assert 0.3 < np.mean(np.log(n) > np.log(8000)) < 0.7  # log-uniform around the geometric midpoint
    stream = serving.served_3d_specs(count=50, rate=2.0)
    gaps = np.diff([s.arrival for s in stream])
    assert np.all(gaps >= 0) and 0.2 < gaps.mean() < 1.0
    assert all(20000 * 0.8 < s.fluid_count() < 2000000 * 1.2 for s in stream)


def test_admission_and_determinism_on_the_oracle(oracle):
    specs = serving.batched_2d_specs(count=4, n_lo=150, n_hi=400, steps=3, seed=11)
    for s in specs:
        s.precision = capi.FP64  # the oracle restates the FP64 path
    alone, _ = serving.Server(oracle, max_active=1).run(specs, keep_fields=True, realtime=False)
    budget = max(s.footprint_bytes() for s in specs) * 2
    mixed, _ = serving.Server(oracle, max_active=4, hbm_budget=budget).run(specs, keep_fields=True,
                                                                          realtime=False)
    spans = sorted((r.start, r.finish) for r in mixed)
    for t in [s for s, _ in spans]:
        assert sum(1 for s, f in spans if s <= t < f) <= 2
    for a, b in zip(alone, mixed):
        assert a.acoustic_steps == b.acoustic_steps > 0
        for k in a.fields:
            assert np.array_equal(a.fields[k], b.fields[k])
