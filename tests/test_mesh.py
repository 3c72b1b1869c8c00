"""Mesh producer and static tables (CPU): invariants that pin the conformal
map, the panel frames, the edge-halo permutation and the metric restatement
independently of their code."""
import math

import numpy as np
import pytest

# panel frames: global = R_p local; local face centre (0,0,1)
R = np.array([[[0, 0, 1], [1, 0, 0], [0, 1, 0]], [[-1, 0, 0], [0, 0, 1], [0, 1, 0]],
              [[0, 0, -1], [-1, 0, 0], [0, 1, 0]], [[1, 0, 0], [0, 0, -1], [0, 1, 0]],
              [[0, -1, 0], [1, 0, 0], [0, 0, 1]], [[0, 1, 0], [1, 0, 0], [0, 0, -1]]], float)


@pytest.fixture(scope="module")
def mesh32(ccs):
    return ccs.Mesh(32, halo=4)


def test_unit_sphere_and_conformal(mesh32):
    m = mesh32
    xyz = m.xyz.reshape(-1, 3)
    assert np.max(np.abs(np.linalg.norm(xyz, axis=1) - 1)) < 1e-15
    J = m.jab_unit.reshape(-1, 2, 3)
    g11, g22, g12 = (J[:, 0] ** 2).sum(1), (J[:, 1] ** 2).sum(1), (J[:, 0] * J[:, 1]).sum(1)
    assert np.max(np.abs(g12) / g11) < 1e-14
    assert np.max(np.abs(g11 - g22) / g11) < 1e-14
    # the Jacobian is tangent to the sphere
    assert np.max(np.abs((J * xyz[:, None, :]).sum(2))) < 1e-14


def test_map_matches_the_defining_integral(ccs):
    # z(w) = (1/kappa) int_0^w (t^8 + 14 t^4 + 1)^(-1/4) dt, kappa = (varpi/2)/108^(1/4)
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 25
    n = 16
    m = ccs.Mesh(n, halo=3)
    loc = m.xyz.reshape(6, n, n, 3)[0] @ R[0]
    w = (loc[..., 0] + 1j * loc[..., 1]) / (1 + loc[..., 2])
    kappa = mp.quad(lambda t: 1 / mp.sqrt(1 - t ** 4), [0, 1]) / mp.mpf(108) ** 0.25
    for j, i in [(8, 8), (15, 15), (15, 8), (0, 3), (11, 14)]:
        wi = mp.mpc(w[j, i])
        z = mp.quad(lambda t: wi * ((wi * t) ** 8 + 14 * (wi * t) ** 4 + 1) ** (-0.25), [0, 0.5, 0.9, 1]) / kappa
        assert abs(complex(z) - complex(-1 + (2 * i + 1) / n, -1 + (2 * j + 1) / n)) < 1e-14


def test_edge_midpoint_and_vertex_closed_forms(ccs):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 30
    # varpi/2 = Gamma(1/4)^2 / (4 sqrt(2 pi))  (lemniscate constant)
    kappa = mp.gamma(mp.mpf(1) / 4) ** 2 / (4 * mp.sqrt(2 * mp.pi)) / mp.mpf(108) ** 0.25
    wE = mp.sqrt(2) - 1                        # tan(pi/8): edge midpoint
    assert abs(mp.quad(lambda t: (t ** 8 + 14 * t ** 4 + 1) ** (-0.25), [0, wE]) / kappa - 1) < 1e-20
    wV = (1 + 1j) / (mp.sqrt(3) + 1)           # cube vertex: zero of W
    assert abs(wV ** 8 + 14 * wV ** 4 + 1) < 1e-20


@pytest.mark.parametrize("h", [3, 4])
def test_halo_permutation_is_the_reflection_across_the_edge(ccs, orc, h):
    """The halo copied by the reference's table must sit where the panel's
    own map, continued across its edge, puts it: the mirror image of the
    panel's cell across the edge's great-circle plane."""
    n = 24
    m = ccs.Mesh(n, halo=h)
    P = n + 2 * h
    comp = []
    for a in range(3):
        f = np.zeros(P * P * 6)
        V = f.reshape(6, P, P)
        V[:, h:h + n, h:h + n] = m.xyz.reshape(6, n, n, 3)[..., a]
        orc.lib().oracle_fill_halo(n, h, f.ctypes.data_as(orc._D))
        comp.append(V.copy())
    X = np.stack(comp, -1)  # [p, j, i, 3] padded
    normals = {"W": (1, 0, 1), "E": (1, 0, -1), "S": (0, 1, 1), "N": (0, 1, -1)}
    worst = 0.0
    for p in range(6):
        for side, nl in normals.items():
            nv = R[p] @ (np.array(nl) / math.sqrt(2))
            for k in range(1, h + 1):
                for t in range(n):
                    if side == "W":
                        halo, mirror = X[p, h + t, h - k], X[p, h + t, h + k - 1]
                    elif side == "E":
                        halo, mirror = X[p, h + t, h + n - 1 + k], X[p, h + t, h + n - k]
                    elif side == "S":
                        halo, mirror = X[p, h - k, h + t], X[p, h + k - 1, h + t]
                    else:
                        halo, mirror = X[p, h + n - 1 + k, h + t], X[p, h + n - k, h + t]
                    refl = mirror - 2 * np.dot(mirror, nv) * nv
                    worst = max(worst, float(np.max(np.abs(refl - halo))))
    assert worst < 1e-14


def test_halo_vector_rotation_is_a_signed_permutation(mesh32):
    """W = (G A^-1)_halo (A G^-1)_source maps covariant components of the
    source panel to the destination panel: a signed permutation to round-off,
    with the pattern of the Jacobian column swaps (ghost_mod.f90:100-153)."""
    m = mesh32
    n, h, P = m.n, m.halo, m.P
    G = m.G.reshape(6, P, P, 2, 2).transpose(0, 1, 2, 4, 3)  # -> [p, j, i, r, c]
    IG = m.IG.reshape(6, P, P, 2, 2).transpose(0, 1, 2, 4, 3)
    A = m.A.reshape(6, P, P, 2, 2).transpose(0, 1, 2, 4, 3)
    IA = m.IA.reshape(6, P, P, 2, 2).transpose(0, 1, 2, 4, 3)
    lon = m.view(m.lon)
    lat = m.view(m.lat)
    worst = 0.0
    for p in range(6):
        for (jj, ii) in [(h + 5, h - 1), (h + 5, h + n), (h - 1, h + 7), (h + n, h + 7)]:
            # source cell: same geographic point in the interior of some panel
            d = (np.abs(m.view(m.lon)[:, h:h + n, h:h + n] - lon[p, jj, ii]) +
                 np.abs(m.view(m.lat)[:, h:h + n, h:h + n] - lat[p, jj, ii]))
            q, js, is_ = np.unravel_index(np.argmin(d), d.shape)
            W = (G[p, jj, ii] @ IA[p, jj, ii]) @ (A[q, js + h, is_ + h] @ IG[q, js + h, is_ + h])
            Wr = np.round(W)
            assert set(np.abs(Wr).sum(0)) == {1.0} and set(np.abs(Wr).sum(1)) == {1.0}
            worst = max(worst, float(np.max(np.abs(W - Wr))))
    assert worst < 1e-12


def test_metric_identities(mesh32):
    m = mesh32
    P = m.P
    mask = ~m.corner_mask().reshape(-1)
    G = m.G.reshape(-1, 2, 2).transpose(0, 2, 1)[mask]
    IG = m.IG.reshape(-1, 2, 2).transpose(0, 2, 1)[mask]
    A = m.A.reshape(-1, 2, 2).transpose(0, 2, 1)[mask]
    IA = m.IA.reshape(-1, 2, 2).transpose(0, 2, 1)[mask]
    I2 = np.eye(2)
    assert np.max(np.abs(G @ IG - I2)) < 1e-13
    assert np.max(np.abs(A @ IA - I2)) < 1e-12
    assert np.max(np.abs((G @ IA) @ (A @ IG) - I2)) < 1e-12
    sg = m.sqrtG.reshape(-1)[mask]
    assert np.max(np.abs(np.sqrt(np.linalg.det(G)) - sg) / sg) < 1e-14
    # conformal: G = sqrtG I
    assert np.max(np.abs(G[:, 0, 1]) / sg) < 1e-13
    assert np.max(np.abs(G[:, 0, 0] / sg - 1)) < 1e-13
    assert P == m.n + 2 * m.halo


def test_sphere_area_converges(ccs):
    errs = []
    for n in (16, 32, 64):
        m = ccs.Mesh(n, halo=3)
        area = float(np.sum(m.interior(m.areaCell)))
        errs.append(abs(area / (4 * math.pi * m.radius ** 2) - 1))
    assert errs[2] < 1e-5
    assert errs[0] / errs[1] > 4 and errs[1] / errs[2] > 4


def test_coriolis_and_latitudes(mesh32):
    m = mesh32
    f = m.interior(m.f)
    lat = m.interior(m.lat)
    assert np.max(np.abs(f - 2 * m.omega * np.sin(lat))) < 1e-18
    assert abs(np.max(lat) - math.pi / 2) < 0.05 and abs(np.min(lat) + math.pi / 2) < 0.05


def test_rejects_bad_sizes(ccs):
    with pytest.raises(ccs.CcsError):
        ccs.Mesh(7, halo=3)   # odd n: no cell centre may sit on a pole
    with pytest.raises(ccs.CcsError):
        ccs.Mesh(2, halo=3)


def test_product_metrics_match_oracle_restatement(ccs, orc):
    for h in (3, 4):
        m = ccs.Mesh(20, halo=h)
        o = orc.metrics(m.n, m.halo, m.radius, m.omega, m.lon_deg, m.lat_deg, m.jab_unit)
        assert o["status"] == 0 and o["dx"] == m.dx
        for k in ("lon", "lat", "sinlat", "coslat", "sqrtG", "f", "areaCell", "G", "IG", "A", "IA"):
            assert np.array_equal(getattr(m, k), o[k]), k


def test_extended_lattice_metric(ccs, mesh32):
    # mesh%matrixIG_ext (io_mod.F90:223-250) on the (2n+1)^2 lattice
    n, m = 32, 65
    ig = mesh32.ig_ext().reshape(6, m, m, 2, 2)  # [p, b, a, c, r]
    corners = np.zeros((m, m), bool)
    corners[[0, 0, -1, -1], [0, -1, 0, -1]] = True
    # the four corner points per panel are never set by the reference
    assert np.all(ig[:, corners] == 0.0)
    ok = ig[:, ~corners]
    # symmetric and conformal: G^-1 = I / sqrtG (unit sphere scaled by a^2)
    assert np.array_equal(ok[..., 0, 1], ok[..., 1, 0])
    assert np.max(np.abs(ok[..., 0, 1] * mesh32.radius ** 2)) < 1e-9
    assert np.max(np.abs(ok[..., 0, 0] / ok[..., 1, 1] - 1.0)) < 1e-9
    # at even lattice indices (cell centres) it is the cell metric matrixIG
    # (the two conformal-map solves agree to round-off)
    P, h = mesh32.P, mesh32.halo
    cell = mesh32.IG.reshape(6, P, P, 2, 2)[:, h:h + n, h:h + n]
    cen = ig[:, 1::2, 1::2]
    assert np.max(np.abs(cen - cell)) / np.max(np.abs(cell)) < 1e-10


def test_extended_lattice_metric_from_mesh_file(ccs, tmp_path):
    path = str(tmp_path / "mesh_C16.nc")
    ccs.write_mesh_file(path, 16)
    n, jab_ext = ccs.read_mesh_file_ext(path)
    assert n == 16 and jab_ext.size == 36 * 33 * 33
    mesh = ccs.Mesh(16, halo=3)
    # the file holds this library's own map values, so the tables agree exactly
    assert np.array_equal(mesh.ig_ext(jab_ext), mesh.ig_ext())
