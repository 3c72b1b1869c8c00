"""The drop-in FP64 GPU path against the reference's own CPU implementation.

The C180 configuration's bar: after one simulated day on identical seeded
inputs (a = 6.371e6 m, balanced jet + 8 bumps, flat bottom, SSP-RK3 at CFL
0.5), normalised L2(h) <= 1e-12 between the GPU and the reference CPU
code.  The reference side is oracle/_ref/libccs_ref_omp.so, its
dynamic-core modules compiled from source with its own GNU flags (OpenMP
on), run by tests/ref_run.py in a child process on all host cores.  Skipped
where that library was not built.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(ROOT, "oracle", "_ref", "libccs_ref_omp.so")


def _cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


@pytest.mark.skipif(not os.path.exists(LIB), reason="the reference's OpenMP build (oracle/_ref) is not built")
def test_one_day_c180_matches_reference_cpu_code(ccs, tmp_path):
    n = 180
    mesh = ccs.Mesh(n, halo=3, radius=6.371e6, omega=7.292e-5)
    s0 = ccs.init_case(mesh, case_num=7, seed=1, n_bumps=8, mountain=False)
    dt, nsteps = ccs.cfl_time_step(mesh, s0, 0.5, 86400.0)
    print(f"1 day C{n}: {nsteps} steps of dt={dt:.3f}; reference on {_cores()} host threads", flush=True)
    out = tmp_path / "ref.npz"
    env = dict(os.environ, OMP_NUM_THREADS=str(min(_cores(), 16)))
    r = subprocess.run([sys.executable, os.path.join(HERE, "ref_run.py"), str(n), str(nsteps), str(out)],
                       capture_output=True, text=True, env=env, timeout=1500)
    assert r.returncode == 0, r.stderr[-2000:]
    ref = np.load(out)
    assert float(ref["dt"]) == dt and int(ref["nsteps"]) == nsteps
    with ccs.DynamicalCore(mesh, dt) as core:
        core.set_state(s0)
        core.step(nsteps)
        s = core.get_state()
    area = mesh.interior(mesh.areaCell)
    g = 9.80616
    l1, l2, linf = ccs.error_norms(mesh.interior(s.phi) / g, mesh.interior(ref["phi"]) / g, area)
    mask = ~mesh.corner_mask().ravel()
    same = all(np.array_equal(getattr(s, k)[mask], ref[k][mask]) for k in ("phiG", "u", "v", "phi"))
    print(f"  GPU vs reference: normalised L1/L2/Linf(h) = {l1:.3e} {l2:.3e} {linf:.3e}; bitwise equal: {same}",
          flush=True)
    assert l2 <= 1e-12
