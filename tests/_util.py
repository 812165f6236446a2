"""Test helpers: oracle loading, fixture paths, a banded DTW distance."""
import os
import subprocess

import numpy as np

from sphinxsys_amd import recording
from sphinxsys_amd._capi import SphxLibrary

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "build", "libsphx_oracle.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def oracle_lib():
    """The CPU restatement (checker only); built from oracle/ when missing."""
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    lib = SphxLibrary(ORACLE_SO)
    assert lib.backend == "oracle-cpu"
    lib.lib.sphx_oracle_set_threads.restype = int
    if not os.environ.get("SPHX_ORACLE_THREADS"):
        # the CPUs this process may run on (a GPU host's share of its cores), at most 16
        lib.lib.sphx_oracle_set_threads(min(16, len(os.sched_getaffinity(0))))
    return lib


# the regression files and the reference's DTW rule live in the package
# (sphinxsys_amd.recording); the tests use the same functions
read_result_xml = recording.read_result_xml
read_dtw_threshold = recording.read_dtw_xml
dtw_distance = recording.dtw_distance


def dtw_distance_banded(a, b, window=5):
    """The same recurrence with the cells outside the band at +infinity, so that
    the end cell is always reached through the band.  The reference's table
    starts at zero everywhere, which makes its distance 0 whenever the second
    series is longer than the first by the band width or more."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    la, lb = len(a), len(b)
    window = max(window, abs(la - lb) + 1)
    d = np.full((la, lb), np.inf)
    d[0, 0] = abs(a[0] - b[0])
    for i in range(1, la):
        d[i, 0] = d[i - 1, 0] + abs(a[i] - b[0])
    for j in range(1, lb):
        d[0, j] = d[0, j - 1] + abs(a[0] - b[j])
    for i in range(1, la):
        lo, hi = max(1, i - window), min(lb, i + window)
        for j in range(lo, hi):
            d[i, j] = abs(a[i] - b[j]) + min(d[i - 1, j], d[i, j - 1], d[i - 1, j - 1])
    return d[la - 1, lb - 1]
