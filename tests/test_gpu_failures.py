"""Failure handling, forced from the host (every kernel still runs to its end:
NaN arithmetic does not fault, and no address depends on the data).

The reference stops the process on bad input (io_mod.F90:604-612 pattern);
here the C ABI returns CCS_ERR_NONFINITE and the handle stays usable:
- ccs_set_state rejects a state that holds NaN or Inf;
- a state that blows up (negative depth makes the Rusanov speed sqrt(g h)
  NaN) is reported by ccs_diagnostics, and a fresh ccs_set_state recovers.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NONFINITE = 4


def _case(ccs):
    mesh = ccs.Mesh(24, halo=3)
    s0 = ccs.init_case(mesh, case_num=2)
    dt, _ = ccs.cfl_time_step(mesh, s0, 0.5, 3600.0)
    return mesh, s0, dt


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp64_fast"])
def test_set_state_rejects_nonfinite(ccs, precision):
    mesh, s0, dt = _case(ccs)
    print(f"{precision}: NaN and Inf in the uploaded state", flush=True)
    with ccs.DynamicalCore(mesh, dt, precision=precision) as core:
        for field, value in (("u", np.nan), ("phiG", np.inf), ("v", -np.inf)):
            bad = s0.copy()
            getattr(bad, field)[mesh.npad // 2] = value
            with pytest.raises(ccs.CcsError) as ei:
                core.set_state(bad)
            assert ei.value.code == NONFINITE
        core.set_state(s0)
        core.step(2)
        m, e = core.diagnostics()
        assert np.isfinite(m) and np.isfinite(e)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_blown_up_state_reports_nonfinite_diagnostics(ccs, precision):
    mesh, s0, dt = _case(ccs)
    bad = s0.copy()
    h = mesh.halo
    mesh.view(bad.phiG)[0, h + 5:h + 8, h + 5:h + 8] *= -1.0  # negative depth on a 3x3 block
    print(f"{precision}: negative depth, 2 steps", flush=True)
    with ccs.DynamicalCore(mesh, dt, precision=precision) as core:
        core.set_state(bad, with_companions=False)
        core.step(2)
        with pytest.raises(ccs.CcsError) as ei:
            core.diagnostics()
        assert ei.value.code == NONFINITE
        assert not np.all(np.isfinite(core.get_state().phiG))
        # a fresh upload recovers the handle
        core.set_state(s0)
        core.step(2)
        m, _ = core.diagnostics()
        assert np.isfinite(m)
