"""A diverged particle stops the advection step with SPHX_E_NONFINITE.

The time-step reductions take +inf for a particle whose velocity, force or
compression is not finite, so the step becomes zero and the single-call
driver returns status 4 before the position update (the reference's max
reductions would drop a NaN instead).  Every kernel still runs to its end:
after the bad value is replaced, the next advection step runs normally on the
same objects.  The oracle and the HIP library follow the same rule."""
import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import cases
from sphinxsys_amd._capi import SphxError


def _scene(lib, precision):
    sc = cases.DamBreak(lib, cases.dambreak_2d(dp=0.05), precision)
    sc.initialize()
    sc.advection_step()
    return sc


def _poison_and_recover(sc, field, value):
    good = sc.fluid.read(field)
    bad = good.copy()
    bad[17, 0] = value
    sc.fluid.write(field, bad)
    with pytest.raises(SphxError) as err:
        sc.advection_step()
    assert "status 4" in str(err.value), err.value
    assert "non-finite" in str(err.value)
    with pytest.raises(SphxError):
        sc.acoustic_dt.exec()
    sc.fluid.write(field, good)
    assert sc.advection_step() > 0


@pytest.mark.parametrize("field,value", [("Velocity", np.nan), ("ForcePrior", np.inf)])
def test_oracle_reports_nonfinite_state(oracle, field, value):
    sc = _scene(oracle, capi.FP64)
    _poison_and_recover(sc, field, value)
    sc.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
@pytest.mark.parametrize("field,value", [("Velocity", np.nan), ("ForcePrior", np.inf)])
def test_hip_reports_nonfinite_state(hip, precision, field, value):
    print(f"nonfinite: {field}={value} precision {precision}", flush=True)
    sc = _scene(hip, precision)
    _poison_and_recover(sc, field, value)
    sc.close()
