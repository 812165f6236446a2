"""The neighbour-search + density-summation microbenchmark program at 48^3
(110,592 particles; the 256^3 case is `bench.py --microbench 256`): its
positions (std::mt19937_64 jitter) run through the same stage sequence on
the oracle and the HIP library give bit-identical cell lists and neighbour
lists, FP64 density sums within 1e-12 and FP32-mixed within 1e-6, and the
program's own totals agree with them."""
import json
import os
import subprocess

import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import ck

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def stages(lib, pos, m, precision):
    dp = 1.0 / m
    ctx = ck.Context(lib)
    b = ck.Body(ctx, 3, len(pos), dp, capi.BODY_FLUID, precision)
    b.write("Position", pos)
    b.write("VolumetricMeasure", np.full(len(pos), dp ** 3))
    c = ck.CellLinkedList(b, np.full(3, -4 * dp), np.full(3, 1 + 4 * dp))
    inner = ck.Inner(b, c)
    ck.UpdateCellLinkedList(c).exec()
    ck.ParticleSortCK(b, c).exec()
    ck.UpdateCellLinkedList(c).exec()
    ck.UpdateRelation(inner).exec()
    ck.CompressionSummation(inner).exec()
    out = (c.read(), inner.read(), b.read("CompressionSummation"))
    inner.close(), c.close(), b.close(), ctx.close()
    return out


def test_microbench_program_and_lists(hip, oracle, tmp_path):
    m = 48
    exe = os.path.join(ROOT, "sphinxsys_amd", "lib", "microbench")
    dump = str(tmp_path / "pos.bin")
    out = subprocess.run([exe, "--m", str(m), "--repeat", "1", "--dump", dump], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    res = {r["precision"]: r for r in map(json.loads, out.stdout.splitlines())}
    pos = np.fromfile(dump).reshape(-1, 3)
    assert len(pos) == m ** 3
    dp = 1.0 / m
    lattice = (np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) * dp
    assert np.abs(pos - lattice).max() <= 0.1 * dp
    ref = stages(oracle, pos, m, capi.FP64)
    for prec, key in ((capi.FP64, "fp64"), (capi.MIXED, "fp32-mixed")):
        got = stages(hip, pos, m, prec)
        for g, o in zip(got[0], ref[0]):
            assert np.array_equal(g, o)
        for g, o in zip(got[1], ref[1]):
            assert np.array_equal(g, o)
        err = np.linalg.norm(got[2] - ref[2]) / np.linalg.norm(ref[2])
        assert err < (1e-12 if prec == capi.FP64 else 1e-6), (key, err)
        assert res[key]["neighbours_total"] == len(ref[1][2])
        assert res[key]["summation_mean"] == pytest.approx(ref[2].mean(), rel=1e-6)
