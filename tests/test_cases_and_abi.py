"""Initial conditions and the C ABI surface (CPU only: no call reaches the GPU)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("h", [3, 4])
@pytest.mark.parametrize("case_num,mountain", [(2, False), (5, False), (6, False), (7, False), (7, True)])
def test_product_cases_match_oracle_restatement(ccs, orc, h, case_num, mountain):
    m = ccs.Mesh(16, halo=h)
    s = ccs.init_case(m, case_num=case_num, seed=11, mountain=mountain)
    o = orc.init_case(m, case_num=case_num, seed=11, mountain=mountain)
    for k in ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid", "phis"):
        assert np.array_equal(getattr(s, k), o[k]), k


def test_seeded_recipe_properties(ccs):
    m = ccs.Mesh(48, halo=3, radius=6.371e6)
    area = m.interior(m.areaCell)
    flat = ccs.init_case(m, case_num=7, n_bumps=0, mountain=False)
    # area-mean depth of the balanced jet is h_mean (mean of sin^2 over the sphere is 1/3)
    hbar = np.sum(m.interior(flat.phi) / 9.80616 * area) / np.sum(area)
    assert abs(hbar - 10000.0) < 0.05
    # zonal wind recovered from the covariant components
    assert np.max(np.abs(m.interior(flat.zonal) - 20.0 * np.cos(m.interior(m.lat)))) < 1e-12
    # seeds: same seed -> same state, different seed -> different bumps
    a = ccs.init_case(m, case_num=7, seed=1)
    b = ccs.init_case(m, case_num=7, seed=1)
    c = ccs.init_case(m, case_num=7, seed=2)
    assert np.array_equal(a.phi, b.phi) and not np.array_equal(a.phi, c.phi)
    # bumps stay within +-100 m summed over 8 bumps
    dh = (m.interior(a.phi) - m.interior(flat.phi)) / 9.80616
    assert np.max(np.abs(dh)) < 800 and np.max(np.abs(dh)) > 1
    # mountain: 2000 m peak carried in phis, free surface kept by phi = phi_bal - g hs
    mt = ccs.init_case(m, case_num=7, seed=1, mountain=True)
    hs = m.interior(mt.phis) / 9.80616
    assert 1900 < hs.max() <= 2000.0 + 1e-9
    assert np.max(np.abs((m.interior(mt.phi) + m.interior(mt.phis)) - m.interior(a.phi))) < 1e-9


def test_case2_is_in_gradient_wind_balance_form(ccs):
    # Williamson case 2 (test_case_mod.F90:61-95): gh = gh0 - (a Omega u0 + u0^2/2) sin^2(lat)
    m = ccs.Mesh(16, halo=3)
    s = ccs.init_case(m, case_num=2)
    u0 = 2 * math.pi * m.radius / (12 * 86400)
    lat = m.interior(m.lat)
    expect = 29400 - (m.radius * m.omega * u0 + u0 ** 2 / 2) * np.sin(lat) ** 2
    assert np.max(np.abs(m.interior(s.phi) - expect)) < 1e-8


def test_cfl_time_step_divides_the_run(ccs):
    m = ccs.Mesh(32, halo=3)
    s = ccs.init_case(m, case_num=7)
    dt, nsteps = ccs.cfl_time_step(m, s, 0.5, 86400.0)
    assert abs(dt * nsteps - 86400.0) < 1e-6
    dt2, _ = ccs.cfl_time_step(m, s, 0.25, 86400.0)
    assert dt2 < dt


def test_error_norms_definitions(ccs):
    r = np.array([1.0, 2.0, 3.0])
    x = np.array([1.0, 2.5, 3.0])
    a = np.array([1.0, 2.0, 1.0])
    l1, l2, linf = ccs.error_norms(x, r, a)
    assert math.isclose(l1, 1.0 / 8.0)
    assert math.isclose(l2, math.sqrt(0.5 / 18.0))
    assert math.isclose(linf, 1.0 / 4.0)


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "ccs.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ccs_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from conformal_cubed_sphere_amd import _lib

    names = _declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(_lib.lib, name), name
        assert name in _lib.SIGNATURES, name


def test_version_string():
    from conformal_cubed_sphere_amd import _lib

    assert b"gfx950" in _lib.lib.ccs_version()


def test_create_rejects_invalid_configs_before_touching_a_device():
    from conformal_cubed_sphere_amd import _lib

    h = ctypes.c_void_p()
    for bad in (dict(n=3), dict(halo=2), dict(precision=9), dict(scheme=5), dict(dt=0.0), dict(eflux=1),
                dict(spatial_op=5), dict(spatial_op=-1),
                # the extended-lattice operator runs on the FP64 parity path only
                dict(spatial_op=4, precision=1), dict(spatial_op=4, precision=2)):
        cfg = _lib.CcsConfig(n=32, halo=3, precision=0, scheme=0, eflux=0, device=0, use_graph=0, dt=60.0,
                             dx=0.05, radius=6.371e6, gravity=9.80616, omega=7.292e-5)
        for k, v in bad.items():
            setattr(cfg, k, v)
        assert _lib.lib.ccs_create(ctypes.byref(cfg), ctypes.byref(h)) == _lib.CCS_ERR_INVALID
    assert _lib.lib.ccs_create(None, ctypes.byref(h)) == _lib.CCS_ERR_INVALID


def test_null_handle_calls_fail_cleanly():
    from conformal_cubed_sphere_amd import _lib

    L = _lib.lib
    assert L.ccs_step(None, 1) == _lib.CCS_ERR_INVALID
    assert L.ccs_synchronize(None) == _lib.CCS_ERR_INVALID
    assert L.ccs_last_error(None) == b"null handle"
    L.ccs_destroy(None)


def test_mesh_metrics_rejects_bad_arguments():
    from conformal_cubed_sphere_amd import _lib

    m = _lib.CcsMetrics()
    m.n, m.halo = 16, 0
    x = np.zeros(16 * 16 * 6)
    assert _lib.lib.ccs_mesh_metrics(_lib.dptr(x), _lib.dptr(x), _lib.dptr(np.zeros(6 * x.size)),
                                     ctypes.byref(m)) == _lib.CCS_ERR_INVALID
    assert _lib.lib.ccs_mesh_geometry(5, None, None, None, None) == _lib.CCS_ERR_INVALID


def test_ensemble_create_rejects_invalid_configs_before_touching_a_device():
    from conformal_cubed_sphere_amd import _lib

    h = ctypes.c_void_p()
    n = (ctypes.c_int * 2)(32, 48)
    dt = np.array([60.0, 40.0])
    dx = np.array([0.05, 0.03])
    for bad in (dict(n_members=0), dict(halo=2), dict(precision=0), dict(scheme=3), dict(scheme=-1), dict(tile_y=12),
                dict(persistent=2)):
        cfg = _lib.CcsEnsembleConfig(n_members=2, halo=3, precision=1, scheme=0, device=0, use_graph=1, tile_y=0,
                                     persistent=-1, radius=6.371e6, gravity=9.80616, omega=7.292e-5)
        for k, v in bad.items():
            setattr(cfg, k, v)
        assert _lib.lib.ccs_ensemble_create(ctypes.byref(cfg), n, _lib.dptr(dt), _lib.dptr(dx),
                                            ctypes.byref(h)) == _lib.CCS_ERR_INVALID, bad
    cfg = _lib.CcsEnsembleConfig(n_members=2, halo=3, precision=1, scheme=0, device=0, use_graph=1, tile_y=0,
                                 persistent=-1, radius=6.371e6, gravity=9.80616, omega=7.292e-5)
    bad_dt = np.array([60.0, 0.0])
    assert _lib.lib.ccs_ensemble_create(ctypes.byref(cfg), n, _lib.dptr(bad_dt), _lib.dptr(dx),
                                        ctypes.byref(h)) == _lib.CCS_ERR_INVALID
    assert _lib.lib.ccs_ensemble_step(None, 1) == _lib.CCS_ERR_INVALID
    assert _lib.lib.ccs_ensemble_last_error(None) == b"null ensemble"
    _lib.lib.ccs_ensemble_destroy(None)


def test_set_mesh_ext_rejects_null_and_handles():
    from conformal_cubed_sphere_amd import _lib

    assert _lib.lib.ccs_set_mesh_ext(None, None) == _lib.CCS_ERR_INVALID
    assert _lib.lib.ccs_mesh_ext_metric(5, 1.0, None, _lib.dptr(np.zeros(4))) == _lib.CCS_ERR_INVALID
    assert _lib.lib.ccs_mesh_ext_metric(8, 1.0, None, None) == _lib.CCS_ERR_INVALID
