"""Observation recording and the regression harness's file formats.

Restates the reference's output side of an observation so that series produced
on the device land in the files its regression tooling reads:

  ObservedQuantityRecording   shared_ck/io_system/io_observation_ck.h:39-104
      <Observer>_<Quantity>.dat: header `run_time   "Pressure[0]"   ...`, then per
      write the physical time (default stream format) and every observed value
      (fixed, 9 decimals), fields separated by three spaces (io_plt.cpp:13-34)
  ReducedQuantityRecording    io_observation_ck.h:106-158
      <Body>_<Quantity>.dat: header `"run_time"   "<Quantity>"   `
  RegressionTestDynamicTimeWarping
                              regression_test/regression_test_base.{h,hpp},
                              dynamic_time_warping_method.{h,hpp}
      observations remembered per write; <id>_<Quantity>_Run_<n>_result.xml
      (Snapshot_Element / Result_Element, Particle_k with snapshot_l
      attributes), <id>_<Quantity>_dtwdistance.xml, <id>_<Quantity>_runtimes.dat,
      the DTW distance and the test / database-generation procedures.

XML attribute values follow the third-party Simbody SimTKcommon `String(value)`
conversion the reference's XmlEngine uses (xml_engine.h:68-72): "%.17g" for a
double build, "%.9g" for a float build (`real_type`), NaN and infinities as
"NaN", "Inf", "-Inf".  The reference's archived result and runtimes files are
reproduced byte for byte by these writers (some archived DTW thresholds are
round values set by hand, which no build writes).
"""
import math
import os
import xml.etree.ElementTree as ET

import numpy as np

XML_HEAD = '<?xml version="1.0" encoding="UTF-8" ?>\n'


# --------------------------------------------------------------------------
# value formatting
# --------------------------------------------------------------------------
def format_real(x, real_type=np.float64):
    """SimTK::String of a Real: %.17g (double) or %.9g (float)."""
    x = float(np.float32(x)) if real_type == np.float32 else float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    return ("%.9g" if real_type == np.float32 else "%.17g") % x


def format_value(v, real_type=np.float64):
    """SimTK::String of a Real (format_real) or of a Vec: "~[" components "," "]"
    in the stream's default format (%g, precision 6)."""
    a = np.asarray(v, dtype=np.float64)
    if a.ndim == 0:
        return format_real(float(a), real_type)
    return "~[" + ",".join("%g" % float(x) for x in a.reshape(-1)) + "]"


def parse_value(text):
    """A Real, or a Vec written as "~[a,b,...]" (returned as an array)."""
    if text.startswith("~["):
        return np.array([float(x) for x in text[2:-1].split(",")])
    return float(text)


def format_stream_default(x):
    """`std::ostream << double` with the default format (%g, precision 6)."""
    return "%g" % float(x)


def format_fixed9(x):
    """`std::fixed << std::setprecision(9)` (io_plt.cpp:31-34)."""
    return "%.9f" % float(x)


# --------------------------------------------------------------------------
# .dat recordings
# --------------------------------------------------------------------------
class ObservedQuantityRecording:
    """Interpolated quantity at the observer points written to
    `<observer>_<quantity>.dat` in `output_folder` on every write_to_file; a
    vector quantity (`dim` components) writes "<quantity>[i][k]" columns
    (io_plt.hpp:9-31)."""

    def __init__(self, observe, n_observe, quantity_name, observer_name, output_folder, dim=None):
        self.observe = observe            # callable returning the observed values
        self.n = int(n_observe)
        self.dim = dim
        self.quantity_name = quantity_name
        self.identifier = observer_name
        self.path = os.path.join(output_folder, f"{observer_name}_{quantity_name}.dat")
        self.header_written = False
        self.last = None

    def write_to_file(self, physical_time, iteration=0):
        if not self.header_written:
            with open(self.path, "w") as f:
                f.write("run_time" + "   ")
                for i in range(self.n):
                    name_i = f"{self.quantity_name}[{i}]"
                    if self.dim is None:
                        f.write('"' + name_i + '"' + "   ")
                    else:
                        for k in range(self.dim):
                            f.write('"' + name_i + "[" + str(k) + ']"' + "   ")
                f.write("\n")
            self.header_written = True
        shape = (self.n,) if self.dim is None else (self.n, self.dim)
        values = np.asarray(self.observe(), dtype=np.float64).reshape(shape)
        with open(self.path, "a") as f:
            f.write(format_stream_default(physical_time) + "   ")
            for v in values.reshape(-1):
                f.write(format_fixed9(v) + "   ")
            f.write("\n")
        self.last = values
        return values

    def observed(self):
        return self.last


class ReducedQuantityRecording:
    """A reduced quantity of a body written to `<body>_<quantity>.dat`."""

    def __init__(self, reduce, quantity_name, body_name, output_folder):
        self.reduce = reduce              # callable returning one value
        self.n = 1
        self.quantity_name = quantity_name
        self.identifier = body_name
        self.path = os.path.join(output_folder, f"{body_name}_{quantity_name}.dat")
        self.header_written = False
        self.last = None

    def write_to_file(self, physical_time, iteration=0):
        if not self.header_written:
            with open(self.path, "w") as f:
                f.write('"run_time"' + "   " + '"' + self.quantity_name + '"' + "   " + "\n")
            self.header_written = True
        value = float(self.reduce())
        with open(self.path, "a") as f:
            f.write(format_stream_default(physical_time) + "   " + format_fixed9(value) + "   " + "\n")
        self.last = np.array([value])
        return value

    def observed(self):
        return self.last


# --------------------------------------------------------------------------
# regression XML files
# --------------------------------------------------------------------------
def write_result_xml(path, series, real_type=np.float64):
    """`*_Run_<n>_result.xml`: series[k][l] is observation k at snapshot l
    (regression_test_base.hpp:207-232)."""
    series = [np.asarray(s, dtype=np.float64) for s in series]
    snapshots = len(series[0]) if series else 0
    out = [XML_HEAD, "<result>\n", "    <Snapshot_Element>\n",
           f'        <Snapshot number_of_snapshot_for_local_result_="{snapshots}" />\n',
           "    </Snapshot_Element>\n", "    <Result_Element>\n"]
    for k, s in enumerate(series):
        attrs = " ".join(f'snapshot_{l}="{format_value(v, real_type)}"' for l, v in enumerate(s))
        out.append(f"        <Particle_{k} {attrs} />\n")
    out += ["    </Result_Element>\n", "</result>\n"]
    with open(path, "w") as f:
        f.write("".join(out))


def read_result_xml(path):
    """Observation series of a `*_result.xml` (one array per Particle_k)."""
    root = ET.parse(path).getroot()
    snapshots = int(root.find("Snapshot_Element")[0].attrib["number_of_snapshot_for_local_result_"])
    series = []
    for el in root.find("Result_Element"):
        series.append(np.array([parse_value(el.attrib[f"snapshot_{l}"]) for l in range(snapshots)]))
    return series


def write_dtw_xml(path, quantity_name, distances, real_type=np.float64):
    """`*_dtwdistance.xml` (dynamic_time_warping_method.hpp:104-117)."""
    attrs = " ".join(f'{quantity_name}_{k}="{format_real(d, real_type)}"' for k, d in enumerate(distances))
    with open(path, "w") as f:
        f.write(XML_HEAD + "<dtw_distance>\n" + f"    <DTWDistance {attrs} />\n" + "</dtw_distance>\n")


def read_dtw_xml(path):
    el = ET.parse(path).getroot().find("DTWDistance")
    return [float(v) for v in el.attrib.values()]


def write_runtimes(path, converged, number_of_run, label_for_repeat):
    """`*_runtimes.dat`: converged flag, number of runs, repeat label
    (regression_test_base.hpp:240-256)."""
    with open(path, "w") as f:
        f.write(f"{'true' if converged else 'false'}\n{int(number_of_run)}\n{int(label_for_repeat)}")


def read_runtimes(path):
    with open(path) as f:
        words = f.read().split()
    return words[0] == "true", int(words[1]), int(words[2])


# --------------------------------------------------------------------------
# DTW distance and the regression procedures
# --------------------------------------------------------------------------
def dtw_distance(a, b, window=5):
    """RegressionTestDynamicTimeWarping::calculateDTWDistance for one
    observation (dynamic_time_warping_method.hpp:16-55): |a_i - b_j| local cost,
    band of half-width max(5, |len(a) - len(b)|), cells outside the band left at 0."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    la, lb = len(a), len(b)
    if lb > 1.2 * la or lb < 0.8 * la:
        raise ValueError("series lengths differ by more than 20%")

    def pnorm(x, y):
        # calculatePNorm: |a - b| for a Real, (a - b).norm() for a vector (dynamic_time_warping_method.h:55-63)
        diff = x - y
        return np.abs(diff) if a.ndim == 1 else np.sqrt(np.sum(diff * diff, axis=-1))

    d = np.zeros((la, lb))
    d[0, 0] = pnorm(a[0], b[0])
    for i in range(1, la):
        d[i, 0] = d[i - 1, 0] + pnorm(a[i], b[0])
    for j in range(1, lb):
        d[0, j] = d[0, j - 1] + pnorm(a[0], b[j])
    window = max(window, abs(la - lb))
    for i in range(1, la):
        lo, hi = max(1, i - window), min(lb, i + window)
        if lo >= hi:
            continue
        cost = pnorm(a[i], b[lo:hi])
        # the row recurrence has a left-neighbour dependence: walk it in order
        up = d[i - 1, lo:hi]
        diag = d[i - 1, lo - 1:hi - 1]
        prev = d[i, lo - 1]
        row = d[i]
        for t in range(hi - lo):
            v = cost[t] + min(up[t], prev, diag[t])
            row[lo + t] = v
            prev = v
    return float(d[la - 1, lb - 1])


class RegressionTestDynamicTimeWarping:
    """Remembers the observations of a recording at every write and compares
    them with the archived runs in `input_folder`
    (regression_test_base.h:109-119, dynamic_time_warping_method.h:88-133)."""

    def __init__(self, recording, input_folder, real_type=np.float64):
        self.rec = recording
        self.input_folder = input_folder
        self.real_type = real_type
        self.prefix = os.path.join(input_folder, f"{recording.identifier}_{recording.quantity_name}")
        self.current = []            # snapshot-major, as current_result_
        self.last_iteration = None

    def write_to_file(self, physical_time, iteration=0):
        if iteration == self.last_iteration:
            raise RuntimeError("the iteration step is not changed (duplicated observation)")
        self.last_iteration = iteration
        self.rec.write_to_file(physical_time, iteration)
        obs = np.array(self.rec.observed(), dtype=np.float64)
        self.current.append(obs.reshape((self.rec.n,) + obs.shape[1:]) if obs.ndim > 1 else obs.reshape(self.rec.n))

    def observed(self):
        return self.rec.observed()

    def current_series(self):
        """current_result_trans_: observation-major series."""
        return [np.array([snap[k] for snap in self.current]) for k in range(self.rec.n)]

    def result_path(self, n):
        return f"{self.prefix}_Run_{n}_result.xml"

    def test_result(self):
        """testResult: for every archived run, DTW(archived, current) of each
        observation must not exceed 1.01 x the archived distance.  Returns
        [(run, distances, limits)] and raises AssertionError on failure."""
        limits = read_dtw_xml(f"{self.prefix}_dtwdistance.xml")
        _, number_of_run, _ = read_runtimes(f"{self.prefix}_runtimes.dat")
        cur = self.current_series()
        report = []
        for n in range(number_of_run):
            if not os.path.exists(self.result_path(n)):
                continue  # "This result has not been preserved and will not be compared."
            archived = read_result_xml(self.result_path(n))
            dist = [dtw_distance(archived[k], cur[k]) for k in range(self.rec.n)]
            report.append((n, dist, limits))
            bad = [k for k in range(self.rec.n) if dist[k] > 1.01 * limits[k]]
            if bad:
                raise AssertionError(f"{self.rec.quantity_name}: run {n} observations {bad} beyond the DTW limit: "
                                     f"{[dist[k] for k in bad]} > 1.01 x {[limits[k] for k in bad]}")
        return report

    def generate_data_base(self, threshold):
        """generateDataBase: while not converged, the new DTW distances are the
        maxima over the archived runs (and the previous maxima); the current
        result becomes run number_of_run - 1; converged after five consecutive
        runs whose maxima moved by at most `threshold`."""
        runtimes = f"{self.prefix}_runtimes.dat"
        converged, number_of_run, label = (False, 1, 0)
        if os.path.exists(runtimes):
            converged, number_of_run, label = read_runtimes(runtimes)
        if converged:
            return True
        cur = self.current_series()
        n_obs = self.rec.n
        dtw_path = f"{self.prefix}_dtwdistance.xml"
        old = read_dtw_xml(dtw_path) if number_of_run > 1 else [0.0] * n_obs
        new = list(old)
        for n in range(number_of_run - 1):
            archived = read_result_xml(self.result_path(n))
            dist = [dtw_distance(cur[k], archived[k]) for k in range(n_obs)]
            new = [max(dist[k], old[k], new[k]) for k in range(n_obs)]
        write_result_xml(self.result_path(number_of_run - 1), cur, self.real_type)
        write_dtw_xml(dtw_path, self.rec.quantity_name, new, self.real_type)
        if number_of_run > 1:
            if all(abs(old[k] - new[k]) <= threshold for k in range(n_obs)):
                if label == 4:
                    converged = True
                else:
                    label += 1
            else:
                label = 0
        write_runtimes(runtimes, converged, number_of_run + (0 if converged else 1), label)
        return converged
