"""ctypes access to the CPU oracle (oracle/build/libccs_oracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "build", "libccs_oracle.so")

_D = ctypes.POINTER(ctypes.c_double)


class OracleMesh(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("h", ctypes.c_int), ("dx", ctypes.c_double), ("dt", ctypes.c_double),
                ("sqrtG", _D), ("G", _D), ("IG", _D), ("A", _D), ("IA", _D), ("f", _D), ("phis", _D),
                ("op", ctypes.c_int), ("pad_", ctypes.c_int), ("IG_ext", _D)]


# spatial operator variants (oracle_mesh.op)
OPS = {"weno": 0, "centred2": 1, "centred4": 2, "centred6": 3, "weno_ext": 4}


class OracleStat(ctypes.Structure):
    _fields_ = [(k, _D) for k in ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid", "phit")]


def _p(a):
    return None if a is None else a.ctypes.data_as(_D)


def load():
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    lib = ctypes.CDLL(ORACLE_LIB)
    lib.oracle_weno_limiter.restype = ctypes.c_double
    lib.oracle_weno_limiter.argtypes = [_D, ctypes.c_int]
    lib.oracle_d6.restype = ctypes.c_double
    lib.oracle_d6.argtypes = [_D, ctypes.c_double]
    lib.oracle_metrics.restype = ctypes.c_int
    lib.oracle_metrics.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _D, _D, _D,
                                   _D] + [_D] * 11
    lib.oracle_init_case.restype = ctypes.c_int
    lib.oracle_init_case.argtypes = ([ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_int,
                                      ctypes.c_int] + [ctypes.c_double] * 7 + [_D] * 7 + [_D] * 9)
    lib.oracle_spatial_operator.restype = None
    lib.oracle_spatial_operator.argtypes = [ctypes.POINTER(OracleMesh), ctypes.POINTER(OracleStat), _D, _D, _D]
    lib.oracle_correct_bdy_ghost.restype = None
    lib.oracle_correct_bdy_ghost.argtypes = [ctypes.POINTER(OracleMesh), ctypes.POINTER(OracleStat)]
    lib.oracle_fill_halo.restype = None
    lib.oracle_fill_halo.argtypes = [ctypes.c_int, ctypes.c_int, _D]
    lib.oracle_steps.restype = ctypes.c_int
    lib.oracle_steps.argtypes = [ctypes.POINTER(OracleMesh), ctypes.POINTER(OracleStat), ctypes.c_int,
                                 ctypes.c_int]
    lib.oracle_diagnostics.restype = None
    lib.oracle_diagnostics.argtypes = [ctypes.POINTER(OracleMesh), ctypes.POINTER(OracleStat), _D, _D, _D, _D]
    lib.oracle_set_threads.restype = None
    lib.oracle_set_threads.argtypes = [ctypes.c_int]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


def weno(q, direction):
    q = np.ascontiguousarray(q, dtype=np.float64)
    return lib().oracle_weno_limiter(_p(q), int(direction))


def d6(f, dh):
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().oracle_d6(_p(f), float(dh))


def metrics(n, h, radius, omega, lon_deg, lat_deg, jab):
    P = n + 2 * h
    npd = P * P * 6
    out = {k: np.zeros(npd) for k in ("lon", "lat", "sinlat", "coslat", "sqrtG", "f", "areaCell")}
    for k in ("G", "IG", "A", "IA"):
        out[k] = np.zeros(4 * npd)
    dx = ctypes.c_double()
    rc = lib().oracle_metrics(n, h, radius, omega, _p(lon_deg), _p(lat_deg), _p(jab), ctypes.byref(dx),
                              *(_p(out[k]) for k in ("lon", "lat", "sinlat", "coslat", "sqrtG", "f", "areaCell",
                                                      "G", "IG", "A", "IA")))
    out["dx"] = dx.value
    out["status"] = rc
    return out


def init_case(mesh, case_num=7, seed=1, n_bumps=8, mountain=False, gravity=9.80616, u0=20.0, h_mean=10000.0,
              bump_amp=100.0, bump_rmin=5.0e5, bump_rmax=1.5e6, mountain_h=2000.0, mountain_r=1.5e6):
    """Oracle restatement of initTestCase on the product mesh's tables."""
    npd = mesh.npad
    a = {k: np.zeros(npd) for k in ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid", "phis")}
    rc = lib().oracle_init_case(mesh.n, mesh.halo, mesh.radius, mesh.omega, gravity, case_num, seed, n_bumps,
                                int(bool(mountain)), u0, h_mean, bump_amp, bump_rmin, bump_rmax, mountain_h,
                                mountain_r, _p(mesh.lon), _p(mesh.lat), _p(mesh.sinlat), _p(mesh.coslat),
                                _p(mesh.sqrtG), _p(mesh.G), _p(mesh.IA),
                                *(_p(a[k]) for k in ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid", "phis")))
    assert rc == 0
    return a


class Oracle:
    """A CPU run of the reference time step on a product Mesh and a State."""

    def __init__(self, mesh, state, dt, threads=0, op="weno", ig_ext=None):
        self.mesh = mesh
        self.phis = np.ascontiguousarray(state.phis, dtype=np.float64)
        if op == "weno_ext" and ig_ext is None:
            ig_ext = mesh.ig_ext()
        self.ig_ext = None if ig_ext is None else np.ascontiguousarray(ig_ext, dtype=np.float64)
        self.m = OracleMesh(mesh.n, mesh.halo, mesh.dx, float(dt), _p(mesh.sqrtG), _p(mesh.G), _p(mesh.IG),
                            _p(mesh.A), _p(mesh.IA), _p(mesh.f), _p(self.phis), OPS[op], 0, _p(self.ig_ext))
        self.a = {}
        for k in ("phiG", "u", "v", "phi", "uc", "vc"):
            self.a[k] = np.ascontiguousarray(getattr(state, k), dtype=np.float64).copy()
        for k in ("zonal", "merid"):
            src = getattr(state, k)
            self.a[k] = np.zeros(mesh.npad) if src is None else src.copy()
        self.s = OracleStat(*(_p(self.a[k]) for k in ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid")),
                            None)
        if threads:
            lib().oracle_set_threads(int(threads))

    def steps(self, nsteps, scheme=0):
        rc = lib().oracle_steps(ctypes.byref(self.m), ctypes.byref(self.s), int(nsteps), int(scheme))
        assert rc == 0

    def tendency(self):
        npd = self.mesh.npad
        t = [np.zeros(npd) for _ in range(3)]
        lib().oracle_spatial_operator(ctypes.byref(self.m), ctypes.byref(self.s), *(_p(x) for x in t))
        return tuple(self.mesh.interior(x).copy() for x in t)

    def diagnostics(self):
        r = [ctypes.c_double() for _ in range(4)]
        lib().oracle_diagnostics(ctypes.byref(self.m), ctypes.byref(self.s), *(ctypes.byref(x) for x in r))
        return tuple(x.value for x in r)
