"""Full 2D dam break (3200 fluid particles, t = 20) on the HIP library in FP64
and FP32-mixed, accepted with the reference's regression rule: DTW distance to
every archived run of the probe pressure and mechanical energy at most 1.01 x
the stored maximum (dynamic_time_warping_method.hpp:164-190; fixtures of
tests/2d_examples/test_2d_dambreak/regression_test_tool)."""
import glob
import os

import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import cases
from tests._util import GOLDEN, dtw_distance, read_dtw_threshold, read_result_xml

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_dambreak_2d_full_run_dtw(hip, precision):
    sc = cases.DamBreak(hip, cases.dambreak_2d(), precision)
    sc.initialize()
    series = {"FluidObserver_Pressure": [], "WaterBody_TotalMechanicalEnergy": []}

    def observe():
        p, e = sc.observe()
        series["FluidObserver_Pressure"].append(p[0])
        series["WaterBody_TotalMechanicalEnergy"].append(e)

    observe()
    while sc.time < 20.0:
        sc.advection_step()
        if sc.iterations % 200 == 0:
            observe()
    sc.close()
    d = os.path.join(GOLDEN, "dambreak_2d")
    for quantity, values in series.items():
        limit = 1.01 * read_dtw_threshold(os.path.join(d, f"{quantity}_dtwdistance.xml"))[0]
        for path in sorted(glob.glob(os.path.join(d, f"{quantity}_Run_*_result.xml"))):
            dist = dtw_distance(read_result_xml(path)[0], values)
            print(f"{quantity} {os.path.basename(path)} precision={precision} dtw={dist:.4f} limit={limit:.4f}")
            assert dist <= limit, (quantity, path, dist)
