"""The batch entry points on the oracle (which advances the scenes of a batch
one after the other): each batched scene equals the same scene advanced with
sphx_advection_step, statuses and tuning report come back, and a diverged
scene is reported without stopping the others."""
import numpy as np

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import cases, ck, serving


def test_oracle_batch_equals_single_scene_steps(oracle):
    specs = serving.batched_2d_specs(3, seed=3, n_lo=200, n_hi=600, steps=3)
    alone = []
    for spec in specs:
        sc = cases.DamBreak(oracle, spec.geometry(), capi.FP64)
        sc.initialize()
        t = 0.0
        for s in range(3):
            t, _ = sc.advection.exec(t, sort_now=(s == 1))
        alone.append((sc.fluid.read("Position"), t))
        sc.close()
    scenes = []
    for spec in specs:
        sc = cases.DamBreak(oracle, spec.geometry(), capi.FP64)
        sc.initialize()
        scenes.append(sc)
    scenes[1].fluid.write("Velocity", np.full((scenes[1].fluid.n, 2), np.nan))
    batch = ck.BatchAdvection(oracle, [sc.advection for sc in scenes])
    for s in range(3):
        st = batch.exec(sort_now=[1 if s == 1 else 0] * 3)
        assert list(st) == [0, 4, 0]
    for k in (0, 2):
        assert np.array_equal(scenes[k].fluid.read("Position"), alone[k][0])
        assert batch.time[k] == alone[k][1]
    rep = batch.tuning()
    assert rep["scenes"] == 3 and rep["steps"] == 3
    batch.close()
    for sc in scenes:
        sc.close()
