"""Corrected lid-driven cavity (2500 fluid particles, FP64, t = 30) on the oracle,
recorded with the reference's observation files and accepted with its own
regression rule: DTW distance of each of the 5 + 5 velocity probes to every
archived run at most 1.01 x the stored threshold
(tests/tests_sycl/2d_examples/test_2d_lid_driven_cavity_corrected_sycl/
lid_driven_cavity_sycl.cpp and its regression_test_tool/*)."""
import os
import shutil

import numpy as np
import pytest

from sphinxsys_amd import cases
from sphinxsys_amd import recording as rec
from tests._util import GOLDEN


@pytest.mark.slow
def test_lid_driven_cavity_against_reference_runs(oracle, tmp_path):
    out = tmp_path / "cavity"
    shutil.copytree(os.path.join(GOLDEN, "lid_driven_cavity"), out)
    case = cases.LidDrivenCavity(oracle, cases.lid_driven_cavity())
    case.initialize()
    recs = []
    for name, interp in (("HorizontalVelocity", case.h_velocity), ("VerticalVelocity", case.v_velocity)):
        obs = rec.ObservedQuantityRecording(interp.exec, 5, "Velocity", name, str(out), dim=2)
        recs.append(rec.RegressionTestDynamicTimeWarping(obs, str(out), np.float32))
    case.run(30.0, 1.0, *recs)
    case.close()
    for r in recs:
        assert len(r.current) == 31
        report = r.test_result()
        assert [run for run, _, _ in report] == ([0, 17, 33] if r.rec.identifier == "HorizontalVelocity"
                                                 else [0, 14, 28])
    # the .dat layout of a vector observation: a header naming each component
    header = open(out / "HorizontalVelocity_Velocity.dat").readline()
    assert header.startswith('run_time   "Velocity[0][0]"')
