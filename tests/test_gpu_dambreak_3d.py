"""3D dam break on the GPU at reduced resolution (dp = 0.02, ~90k fluid
particles) to t = 0.8 s, through impact on the end wall: the FP32-mixed run
against the FP64 run of the same scene.  Gates (the flow is chaotic after
impact, so trajectories are not compared): surge front within 1% at every
observation, time-averaged probe pressure within 2%, impact time within two
observation intervals.  The 8M-particle version of this comparison is
`bench.py --full-run 1.5`."""
import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import cases

pytestmark = pytest.mark.gpu


def run(lib, geom, precision, t_end, dt_obs):
    sc = cases.DamBreak(lib, geom, precision)
    sc.initialize()
    t, p, front = [], [], []
    next_obs = 0.0
    while sc.time < t_end:
        if sc.time >= next_obs:
            pr, _ = sc.observe()
            t.append(sc.time)
            p.append(float(pr[0]))
            front.append(sc.front.exec())
            next_obs += dt_obs
        sc.advection_step()
    sc.close()
    return np.array(t), np.array(p), np.array(front)


def test_dambreak_3d_mixed_against_fp64(hip):
    geom = cases.dambreak_3d(dp=0.02)
    t64, p64, f64 = run(hip, geom, capi.FP64, 0.8, 0.01)
    tmx, pmx, fmx = run(hip, geom, capi.MIXED, 0.8, 0.01)
    n = min(len(f64), len(fmx))
    front_err = np.abs(fmx[:n] - f64[:n]).max() / f64[:n].max()
    print(f"front err {front_err:.2e}; mean probe {pmx[:n].mean():.1f} / {p64[:n].mean():.1f}")
    assert f64[:n].max() > 3.15  # the surge reached the end wall
    assert front_err < 0.01
    assert abs(pmx[:n].mean() - p64[:n].mean()) < 0.02 * abs(p64[:n].mean())
    hit64 = t64[np.argmax(f64 >= 0.998 * f64.max())]
    hitmx = tmx[np.argmax(fmx >= 0.998 * f64.max())]
    assert abs(hit64 - hitmx) <= 0.02 + 1e-9
