"""Host-side helpers of bench.py that build the error and drift fields of
the benchmark line: normalised, area-weighted L1/L2/Linf of h, zonal and
meridional wind (doc/orderCheck.m:55-74) and relative mass/energy drift."""
import numpy as np

import bench

GRAVITY = bench.GRAVITY


def test_field_norms_match_error_norms_per_field(ccs):
    mesh = ccs.Mesh(12, halo=3)
    # case 6 (Rossby-Haurwitz wave): every field, the meridional wind too, is
    # nonzero, so each normalisation is defined
    ref = ccs.init_case(mesh, case_num=6, gravity=GRAVITY)
    st = ccs.perturb_state(mesh, ref, 2.0 ** -20)
    out = bench.field_norms(ccs, mesh, st, ref)
    assert set(out) == {"h", "u", "v"}
    area = mesh.interior(mesh.areaCell)
    pairs = {"h": (st.phi / GRAVITY, ref.phi / GRAVITY), "u": (st.zonal, ref.zonal), "v": (st.merid, ref.merid)}
    for name, (x, r) in pairs.items():
        want = ccs.error_norms(mesh.interior(x), mesh.interior(r), area)
        assert (out[name]["L1"], out[name]["L2"], out[name]["Linf"]) == want
        assert 0.0 < out[name]["L2"] < 1e-4
    same = bench.field_norms(ccs, mesh, ref, ref)
    assert all(v == 0.0 for f in same.values() for v in f.values())


def test_drift_is_relative_to_the_first_diagnostics():
    d = bench.drift((2.0, 8.0), (2.0 + 1e-12, 8.0 - 4e-9))
    assert np.isclose(d["mass"], 5e-13, rtol=1e-3)
    assert np.isclose(d["energy"], -5e-10, rtol=1e-3)
    assert bench.drift((3.0, 5.0), (3.0, 5.0)) == {"mass": 0.0, "energy": 0.0}
