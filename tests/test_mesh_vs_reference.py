"""The conformal map against the reference mesh generator's own.

oracle/_ref/libccs_ref_mesh.so is the generator's conformal-map library
(mesh_generation/src/cuco.f90 with Purser's pcoco/pmat/pfft/pietc/pkind
modules) compiled unmodified from its sources by `make -C oracle ref`, with
the harness oracle/ref_mesh_harness.F90 around xmtoxc, called as
generateConformalCubedSphere does for gen_method = 'Rancic'
(mesh_mod.f90:150-166).  This library's map (ccs_mesh_geometry at cell
centres, ccs_mesh_file_write on the (2n+1)^2 extended lattice including the
panel edges and vertices) must give the same points and Jacobians to a few
ulp.  CPU only.
"""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_ref", "libccs_ref_mesh.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="reference mesh library not built")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def _ref_points(xm1, xm2, ipan):
    lib = ctypes.CDLL(LIB)
    xm1 = np.ascontiguousarray(xm1, dtype=np.float64)
    xm2 = np.ascontiguousarray(xm2, dtype=np.float64)
    ipan = np.ascontiguousarray(ipan, dtype=np.int32)
    n = xm1.size
    xc, d = np.zeros(3 * n), np.zeros(6 * n)
    lib.ref_mesh_points(ctypes.c_int(n), xm1.ctypes.data_as(_D), xm2.ctypes.data_as(_D), ipan.ctypes.data_as(_I),
                        xc.ctypes.data_as(_D), d.ctypes.data_as(_D))
    # mesh%jab = dxcdxm / M2D / D2R (mesh_mod.f90:163): per radian of map angle
    return xc.reshape(n, 3), d.reshape(n, 2, 3) / 45.0 / (np.pi / 180.0)


@pytest.mark.parametrize("n", [16, 48, 96])
def test_cell_centres_match_reference_generator(ccs, n):
    ni = n * n * 6
    lon, lat, xyz, jab = np.zeros(ni), np.zeros(ni), np.zeros(3 * ni), np.zeros(6 * ni)
    from conformal_cubed_sphere_amd import _lib

    assert _lib.lib.ccs_mesh_geometry(n, _lib.dptr(lon), _lib.dptr(lat), _lib.dptr(xyz), _lib.dptr(jab)) == 0
    c = -1.0 + (2.0 * np.arange(n) + 1.0) / n
    P, J, I = np.meshgrid(np.arange(6), np.arange(n), np.arange(n), indexing="ij")
    xc, d = _ref_points(c[I.ravel()], c[J.ravel()], P.ravel() + 1)
    assert np.max(np.abs(xyz.reshape(-1, 3) - xc)) <= 2e-15
    assert np.max(np.abs(jab.reshape(-1, 2, 3) - d)) <= 5e-14 * np.max(np.abs(d))


@pytest.mark.parametrize("n", [16, 48])
def test_extended_lattice_with_edges_matches_reference_generator(ccs, tmp_path, n):
    from scipy.io import netcdf_file

    path = str(tmp_path / "mesh.nc")
    ccs.write_mesh_file(path, n)
    with netcdf_file(path, "r", mmap=False) as f:
        X = np.stack([f.variables[k][:].astype(np.float64) for k in ("x", "y", "z")], -1).reshape(-1, 3)
        jab = f.variables["jab"][:].astype(np.float64).reshape(-1, 2, 3)
        xi = f.variables["xi"][:].astype(np.float64).ravel()
        eta = f.variables["eta"][:].astype(np.float64).ravel()
    m = 2 * n + 1
    pan = np.repeat(np.arange(1, 7), m * m)
    xc, d = _ref_points(xi, eta, pan)
    err = np.max(np.abs(X - xc), axis=1).reshape(6, m, m)
    assert np.max(err) <= 2e-15, np.unravel_index(np.argmax(err), err.shape)
    assert np.max(np.abs(jab - d)) <= 5e-14 * np.max(np.abs(d))
