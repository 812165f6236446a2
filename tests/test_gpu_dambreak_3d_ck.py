"""The reference's CK 3D dam break (tests/tests_sycl/3d_examples/test_3d_dambreak_sycl/
dambreak.cpp: dp 0.05, 8000 fluid particles, LinearCorrectionMatrix with alpha 0.5,
the corrected Riemann half steps, surface indication, t = 20) on the GPU against the
reference's three archived runs.

Gates:
  * the reference's own regression procedure (testResult through the
    recording classes, which also write FluidObserver_Pressure.dat and
    WaterBody_TotalMechanicalEnergy.dat): DTW(archived, new) of the
    mechanical-energy series (one value per output interval) and of the six
    probe pressure series (one value per advection step) within 1.01 x the
    archived thresholds (dtwdistance.xml);
  * the same dam break physically: every energy value within 0.05 (10 % of the
    initial energy) of each archived run; the probe mean over the last quarter
    of the run within the envelope of the reference's archived runs of this
    case and of the classic 3D dam break of the same tank
    (tests/3d_examples/test_3d_dambreak), widened by 25 %.  The two archived sets disagree with each other by up to a factor
    of 1.8 on that mean at the upper probes (0.07 against 0.13 at y = 0.266).
The archived runs of this case lose less energy in the first seconds than this
implementation of the current reference source (E(1) = 0.5006 archived, 0.5068
here and on the oracle): the compression summation of the current source starts
the interior at the lattice number density 1.0095 (a pressure of 3.8 against a
hydrostatic 1.0).  The first pressure arrival comes 1-6 % later in advection-step
index than in the archived runs (printed, not gated: the archived series carry
no time stamps and the step size follows the flow speed).
The reference's DTW table starts at zero outside its band, so a path along the
band edge skips most of a long series; the banded distances are printed, not
gated."""
import os
import time

import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import cases, recording
from tests._util import GOLDEN, dtw_distance, dtw_distance_banded, read_dtw_threshold, read_result_xml

pytestmark = pytest.mark.gpu

D = os.path.join(GOLDEN, "dambreak_3d_ck")
ENERGY_RUNS = (0, 5, 10)
PRESSURE_RUNS = (0, 10, 20)


def arrival(series, level=0.05):
    return int(np.argmax(np.asarray(series) > level))


def late_mean(series):
    s = np.asarray(series)
    return float(s[-len(s) // 4:].mean())


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_dambreak_3d_ck_against_reference_runs(hip, precision, tmp_path):
    geom = cases.dambreak_3d_ck()
    assert len(geom.fluid_pos) == 8000
    sc = cases.DamBreak(hip, geom, precision, correction=capi.CORRECTION_LINEAR, correction_alpha=0.5,
                        surface_indication=True)
    sc.initialize()
    # the case's recordings: FluidObserver_Pressure.dat / WaterBody_TotalMechanicalEnergy.dat
    # with the regression wrappers over the archived runs (a float build's files)
    out = str(tmp_path)
    p_reg = recording.RegressionTestDynamicTimeWarping(
        recording.ObservedQuantityRecording(sc.probe.exec, 6, "Pressure", "FluidObserver", out), D, np.float32)
    e_reg = recording.RegressionTestDynamicTimeWarping(
        recording.ReducedQuantityRecording(sc.energy.exec, "TotalMechanicalEnergy", "WaterBody", out), D,
        np.float32)
    t0 = time.perf_counter()
    pressure, energy = sc.run_ck_loop(20.0, 1.0, pressure_recording=p_reg, energy_recording=e_reg)
    wall = time.perf_counter() - t0
    print(f"precision={precision} advection steps={sc.iterations} acoustic steps={sc.acoustic_steps} "
          f"wall={wall:.1f}s energy records={len(energy)}", flush=True)
    sc.close()
    assert len(open(os.path.join(out, "FluidObserver_Pressure.dat")).readlines()) == len(pressure) + 1
    # the reference's testResult over every preserved run (raises beyond 1.01 x the archived distance)
    for reg in (e_reg, p_reg):
        for run, dist, limits in reg.test_result():
            print(f"{reg.rec.quantity_name} run {run}: DTW(archived, current) "
                  f"{[round(d, 4) for d in dist]} limits {limits}", flush=True)

    e_refs = [read_result_xml(os.path.join(D, f"WaterBody_TotalMechanicalEnergy_Run_{r}_result.xml"))[0]
              for r in ENERGY_RUNS]
    e_limit = 1.01 * read_dtw_threshold(os.path.join(D, "WaterBody_TotalMechanicalEnergy_dtwdistance.xml"))[0]
    n = min(len(e) for e in e_refs + [energy])
    print("energy " + " ".join(f"{x:.4f}" for x in energy), flush=True)
    for r, ref in zip(ENERGY_RUNS, e_refs):
        d_ref, d_band = dtw_distance(ref, energy), dtw_distance_banded(ref, energy)
        dev = np.abs(energy[:n] - ref[:n]).max()
        print(f"energy run {r}: dtw {d_ref:.4f} (limit {e_limit:.4f}) banded {d_band:.4f}; max dev {dev:.4f}",
              flush=True)
        assert d_ref <= e_limit, r
        assert dev <= 0.05, r

    p_refs = [read_result_xml(os.path.join(D, f"FluidObserver_Pressure_Run_{r}_result.xml"))
              for r in PRESSURE_RUNS]
    classic = [read_result_xml(os.path.join(GOLDEN, "dambreak_3d_classic", f"FluidObserver_Pressure_Run_{r}_result.xml"))
               for r in (0, 9, 18)]
    p_limits = read_dtw_threshold(os.path.join(D, "FluidObserver_Pressure_dtwdistance.xml"))
    for k in range(6):
        for r, ref in zip(PRESSURE_RUNS, p_refs):
            d_ref = dtw_distance(ref[k], pressure[:, k])
            print(f"pressure probe {k} run {r}: dtw {d_ref:.3f} (limit {1.01 * p_limits[k]:.3f}) "
                  f"banded {dtw_distance_banded(ref[k], pressure[:, k]):.1f}", flush=True)
            assert d_ref <= 1.01 * p_limits[k], (r, k)
        arr = [arrival(ref[k]) for ref in p_refs]
        late = [late_mean(ref[k]) for ref in p_refs + classic]
        a, m = arrival(pressure[:, k]), late_mean(pressure[:, k])
        print(f"probe {k}: arrival step {a} (archived {arr}), late mean {m:.4f} (archived, then classic "
              f"{[round(x, 4) for x in late]}), peak {pressure[:, k].max():.2f}", flush=True)
        assert 0.75 * min(late) <= m <= 1.25 * max(late), k
