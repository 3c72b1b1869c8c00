"""Grid sizes at the edges of the tiling: the smallest grids (n = 4, 6, with
halo up to n), a panel one cell wider than a tile (n = 34), and halo 4.

The FP64 parity path must equal the oracle bit for bit.  The FP32 split path
must stay finite and close to the FP64 path, and must conserve mass to FP32
round-off: the face fluxes telescope exactly, but every cell's increment
dt L is rounded to FP32 (at n = 4 the CFL step is large and the drift after
7 steps reaches about 1.3e-9).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(4, 3), (4, 4), (6, 3), (34, 3), (34, 4)]


def _setup(ccs, n, h):
    mesh = ccs.Mesh(n, halo=h)
    s0 = ccs.init_case(mesh, case_num=7, seed=4, mountain=True)
    dt, _ = ccs.cfl_time_step(mesh, s0, 0.5, 6 * 3600.0)
    return mesh, s0, dt


@pytest.mark.parametrize("n,h", CASES)
def test_parity_small_and_ragged_grids(ccs, orc, n, h):
    mesh, s0, dt = _setup(ccs, n, h)
    print(f"FP64 C{n} halo {h}: tendency and 7 RK3 steps", flush=True)
    with ccs.DynamicalCore(mesh, dt) as core:
        core.set_state(s0)
        tg = core.spatial_operator()
        core.step(7)
        s = core.get_state()
    o = orc.Oracle(mesh, s0, dt)
    for a, b in zip(tg, o.tendency()):
        assert np.array_equal(a, b)
    o.steps(7)
    mask = ~mesh.corner_mask().ravel()
    for k in ("phiG", "u", "v", "phi", "uc", "vc"):
        assert np.array_equal(getattr(s, k)[mask], o.a[k][mask]), k


@pytest.mark.parametrize("n,h", CASES)
def test_split_small_and_ragged_grids(ccs, n, h):
    mesh, s0, dt = _setup(ccs, n, h)
    print(f"FP32 C{n} halo {h}: 7 RK3 steps against FP64", flush=True)
    out = {}
    for prec in ("fp32", "fp64"):
        with ccs.DynamicalCore(mesh, dt, precision=prec) as core:
            core.set_state(s0)
            m0, _ = core.diagnostics()
            core.step(7)
            m1, _ = core.diagnostics()
            out[prec] = (core.get_state(), (m1 - m0) / m0)
    s32, drift32 = out["fp32"]
    s64, _ = out["fp64"]
    h32, h64 = mesh.interior(s32.phi), mesh.interior(s64.phi)
    assert np.all(np.isfinite(h32))
    l2 = np.sqrt(np.sum((h32 - h64) ** 2 * mesh.interior(mesh.areaCell)) /
                 np.sum(h64 ** 2 * mesh.interior(mesh.areaCell)))
    print(f"  L2(h) fp32 vs fp64 {l2:.2e}, fp32 mass drift {drift32:.2e}")
    assert l2 < 1e-5
    assert abs(drift32) < 1e-8

