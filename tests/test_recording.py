"""The observation files and the regression harness (sphinxsys_amd.recording)
against the reference's archived files: every archived result, DTW and
runtimes file written by the harness is reproduced byte for byte from its
parsed values (double
build: 2D and classic 3D dam break; float build: CK 3D dam break and the
lid-driven cavity's vector velocities), the .dat
layouts against the stream formats of io_observation_ck.h / io_plt.cpp, the DTW
recurrence against a direct transcription, and the test / database-generation
procedures on the archived runs."""
import os

import numpy as np
import pytest

from sphinxsys_amd import recording as rec
from tests._util import GOLDEN

ARCHIVES = [
    ("dambreak_2d", "FluidObserver_Pressure", (0, 10, 20), np.float64),
    ("dambreak_2d", "WaterBody_TotalMechanicalEnergy", (0, 6, 11), np.float64),
    ("dambreak_3d_ck", "FluidObserver_Pressure", (0, 10, 20), np.float32),
    ("dambreak_3d_ck", "WaterBody_TotalMechanicalEnergy", (0, 5, 10), np.float32),
    ("dambreak_3d_classic", "FluidObserver_Pressure", (0, 9, 18), np.float64),
    ("lid_driven_cavity", "HorizontalVelocity_Velocity", (0, 17, 33), np.float32),
    ("lid_driven_cavity", "VerticalVelocity_Velocity", (0, 14, 28), np.float32),
]


# DTW files whose thresholds are written as the harness writes them; the others
# hold hand-set round values ("0.2", "0.05", "3.0") that neither a double build
# (%.17g: 0.20000000000000001, 3) nor a float build (%.9g: 0.0500000007) writes
HARNESS_WRITTEN_DTW = {("dambreak_2d", "FluidObserver_Pressure"), ("dambreak_3d_ck", "FluidObserver_Pressure")}


@pytest.mark.parametrize("folder,prefix,runs,real", ARCHIVES)
def test_result_xml_round_trip_byte_exact(tmp_path, folder, prefix, runs, real):
    for n in runs:
        src = os.path.join(GOLDEN, folder, f"{prefix}_Run_{n}_result.xml")
        out = tmp_path / "out.xml"
        rec.write_result_xml(out, rec.read_result_xml(src), real)
        assert out.read_bytes() == open(src, "rb").read(), (folder, prefix, n)


@pytest.mark.parametrize("folder,prefix,runs,real", ARCHIVES)
def test_dtw_and_runtimes_round_trip_byte_exact(tmp_path, folder, prefix, runs, real):
    src = os.path.join(GOLDEN, folder, f"{prefix}_dtwdistance.xml")
    quantity = prefix.split("_", 1)[1]
    out = tmp_path / "dtw.xml"
    values = rec.read_dtw_xml(src)
    rec.write_dtw_xml(out, quantity, values, real)
    if (folder, prefix) in HARNESS_WRITTEN_DTW:
        assert out.read_bytes() == open(src, "rb").read()
    else:
        assert rec.read_dtw_xml(out) == pytest.approx(values, rel=1e-7)
    rt = os.path.join(GOLDEN, folder, f"{prefix}_runtimes.dat")
    if os.path.exists(rt):
        out = tmp_path / "runtimes.dat"
        rec.write_runtimes(out, *rec.read_runtimes(rt))
        assert out.read_bytes() == open(rt, "rb").read()


def test_format_real_cases():
    assert rec.format_real(0.0) == "0"
    assert rec.format_real(1.0000000000000016) == "1.0000000000000016"
    assert rec.format_real(0.50000006, np.float32) == "0.50000006"
    assert rec.format_real(2.5) == "2.5"
    assert rec.format_real(float("nan")) == "NaN"
    assert rec.format_real(float("-inf")) == "-Inf"


def test_observed_and_reduced_dat_layout(tmp_path):
    """Header and rows as io_observation_ck.h:66-90 / 127-147 write them: the
    time in the default stream format (%g), values fixed with 9 decimals,
    three-space separators, a trailing separator before the newline."""
    values = iter([np.array([0.0, 1.5, -2.25]), np.array([1e-10, 123.456789123, 7.0])])
    obs = rec.ObservedQuantityRecording(lambda: next(values), 3, "Pressure", "FluidObserver", str(tmp_path))
    obs.write_to_file(0.0)
    obs.write_to_file(0.052345678901)
    text = open(tmp_path / "FluidObserver_Pressure.dat").read()
    assert text == ('run_time   "Pressure[0]"   "Pressure[1]"   "Pressure[2]"   \n'
                    "0   0.000000000   1.500000000   -2.250000000   \n"
                    "0.0523457   0.000000000   123.456789123   7.000000000   \n")
    energy = iter([0.5, 0.49999999951])
    red = rec.ReducedQuantityRecording(lambda: next(energy), "TotalMechanicalEnergy", "WaterBody", str(tmp_path))
    red.write_to_file(0.0)
    red.write_to_file(12.345678)
    text = open(tmp_path / "WaterBody_TotalMechanicalEnergy.dat").read()
    assert text == ('"run_time"   "TotalMechanicalEnergy"   \n'
                    "0   0.500000000   \n"
                    "12.3457   0.500000000   \n")


def dtw_direct(a, b, window=5):
    """dynamic_time_warping_method.hpp:16-55, transcribed loop for loop."""
    la, lb = len(a), len(b)
    d = [[0.0] * lb for _ in range(la)]
    d[0][0] = abs(a[0] - b[0])
    for i in range(1, la):
        d[i][0] = d[i - 1][0] + abs(a[i] - b[0])
    for j in range(1, lb):
        d[0][j] = d[0][j - 1] + abs(a[0] - b[j])
    window = max(window, abs(la - lb))
    for i in range(1, la):
        for j in range(max(1, i - window), min(lb, i + window)):
            d[i][j] = abs(a[i] - b[j]) + min(d[i - 1][j], d[i][j - 1], d[i - 1][j - 1])
    return d[la - 1][lb - 1]


@pytest.mark.parametrize("la,lb", [(40, 40), (40, 44), (44, 40), (40, 47), (60, 50)])
def test_dtw_matches_direct_transcription(la, lb):
    rng = np.random.default_rng(la * 100 + lb)
    a, b = rng.normal(size=la), rng.normal(size=lb)
    assert rec.dtw_distance(a, b) == pytest.approx(dtw_direct(a, b), rel=1e-14, abs=1e-14)


def test_regression_test_on_archived_runs(tmp_path):
    """An archived run tested against the archive passes the reference's rule
    (DTW(archived, current) <= 1.01 x the archived distance); the same run with
    its values raised by 20 % is rejected."""
    src = os.path.join(GOLDEN, "dambreak_3d_ck")
    series = rec.read_result_xml(os.path.join(src, "WaterBody_TotalMechanicalEnergy_Run_5_result.xml"))[0]

    def run(scale):
        it = iter(series * scale)
        red = rec.ReducedQuantityRecording(lambda: next(it), "TotalMechanicalEnergy", "WaterBody", str(tmp_path))
        reg = rec.RegressionTestDynamicTimeWarping(red, src, np.float32)
        for step in range(len(series)):
            reg.write_to_file(0.1 * step, step)
        return reg

    report = run(1.0).test_result()
    assert [n for n, _, _ in report] == [0, 5, 10]
    with pytest.raises(AssertionError):
        run(1.2).test_result()
    with pytest.raises(RuntimeError):
        reg = run(1.0)
        reg.write_to_file(99.0, len(series) - 1)  # same iteration twice


def test_generate_data_base_procedure(tmp_path):
    """Three generation runs: the first writes Run_0 and zero distances, the
    next ones the maxima of DTW(current, archived) over the archived runs;
    the repeat label counts runs whose maxima moved by at most the threshold."""
    base = rec.read_result_xml(os.path.join(GOLDEN, "dambreak_2d", "FluidObserver_Pressure_Run_0_result.xml"))[0]
    currents = [base, base * 1.01, base * 1.011]
    labels = []
    for n, cur in enumerate(currents):
        it = iter(cur)
        obs = rec.ObservedQuantityRecording(lambda: [next(it)], 1, "Pressure", "FluidObserver", str(tmp_path))
        reg = rec.RegressionTestDynamicTimeWarping(obs, str(tmp_path))
        for step in range(len(cur)):
            reg.write_to_file(0.0, step)
        reg.generate_data_base(threshold=0.05)
        converged, runs, label = rec.read_runtimes(tmp_path / "FluidObserver_Pressure_runtimes.dat")
        assert runs == n + 2 and not converged
        labels.append(label)
        assert os.path.exists(tmp_path / f"FluidObserver_Pressure_Run_{n}_result.xml")
    d1 = rec.dtw_distance(currents[1], currents[0])
    d2 = max(rec.dtw_distance(currents[2], currents[0]), rec.dtw_distance(currents[2], currents[1]), d1)
    assert rec.read_dtw_xml(tmp_path / "FluidObserver_Pressure_dtwdistance.xml")[0] == pytest.approx(d2, rel=1e-15)
    assert labels[0] == 0
    assert labels[2] == (labels[1] + 1 if abs(d2 - d1) <= 0.05 else 0)
