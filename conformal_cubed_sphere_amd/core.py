"""Python host of the time-step path: mesh, initial state, the GPU dynamical
core and the reference's error norms.

Arrays follow the reference layout: flat float64 arrays in Fortran order
(i fastest, then j, then panel) over the halo-padded grid, ``(n+2h)^2*6``
values; ``Mesh.view`` reshapes one to ``[panel, j, i]``.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import dptr, lib

GRAVITY = 9.80616        # constants_mod.F90:8
RADIUS_REF = 6371229.0   # constants_mod.F90:11
OMEGA = 7.292e-5         # constants_mod.F90:14

PRECISIONS = {"fp64": _lib.CCS_FP64_PARITY, "fp32": _lib.CCS_FP32_SPLIT, "fp64_fast": _lib.CCS_FP64_FAST}
SCHEMES = {"RK3_TVD": _lib.CCS_RK3_TVD, "RK4": _lib.CCS_RK4}
# spatial_operators_mod variants (enum ccs_spatial_op): the reference's default
# WENO-Z module and the swap-in sources _2nd, _4th, _6th and _WENO
SPATIAL_OPS = {"weno": _lib.CCS_OP_WENO, "centred2": _lib.CCS_OP_CENTRED2, "centred4": _lib.CCS_OP_CENTRED4,
               "centred6": _lib.CCS_OP_CENTRED6, "weno_ext": _lib.CCS_OP_WENO_EXT}


class CcsError(RuntimeError):
    """A libccs entry point returned a nonzero ccs_status."""

    def __init__(self, where: str, code: int, msg: str = ""):
        self.code = code
        name = _lib.STATUS_NAMES.get(code, str(code))
        super().__init__(f"{where}: {name}" + (f" ({msg})" if msg else ""))


def _check(where: str, code: int, handle=None):
    if code != _lib.CCS_OK:
        msg = lib.ccs_last_error(handle).decode() if handle else ""
        raise CcsError(where, code, msg)


class Mesh:
    """Conformal cubed-sphere mesh and the static tables of read_mesh
    (io_mod.F90:15-307): geometry from ccs_mesh_geometry, metrics from
    ccs_mesh_metrics."""

    def __init__(self, n: int, halo: int = 3, radius: float = RADIUS_REF, omega: float = OMEGA):
        self.n, self.halo, self.radius, self.omega = int(n), int(halo), float(radius), float(omega)
        ni = self.n * self.n * 6
        self.lon_deg = np.zeros(ni)
        self.lat_deg = np.zeros(ni)
        self.xyz = np.zeros(3 * ni)
        self.jab_unit = np.zeros(6 * ni)
        _check("ccs_mesh_geometry", lib.ccs_mesh_geometry(
            self.n, dptr(self.lon_deg), dptr(self.lat_deg), dptr(self.xyz), dptr(self.jab_unit)))
        P = self.n + 2 * self.halo
        self.P = P
        npd = P * P * 6
        self.npad = npd
        for name in ("lon", "lat", "sinlon", "coslon", "sinlat", "coslat", "sqrtG", "f", "areaCell"):
            setattr(self, name, np.zeros(npd))
        for name in ("G", "IG", "A", "IA"):
            setattr(self, name, np.zeros(4 * npd))
        m = _lib.CcsMetrics()
        m.n, m.halo, m.radius, m.omega = self.n, self.halo, self.radius, self.omega
        for name in ("lon", "lat", "sinlon", "coslon", "sinlat", "coslat", "sqrtG", "f", "areaCell",
                     "G", "IG", "A", "IA"):
            setattr(m, name, dptr(getattr(self, name)))
        _check("ccs_mesh_metrics", lib.ccs_mesh_metrics(
            dptr(self.lon_deg), dptr(self.lat_deg), dptr(self.jab_unit), ctypes.byref(m)))
        self._metrics = m
        self.dx = m.dx

    # --- layout helpers
    def view(self, a):
        """[panel, j, i] view of a padded scalar field."""
        return a.reshape(6, self.P, self.P)

    def interior(self, a):
        h, n = self.halo, self.n
        return self.view(a)[:, h:h + n, h:h + n]

    def corner_mask(self):
        idx = np.arange(self.P) - self.halo
        out = (idx < 0) | (idx >= self.n)
        return np.broadcast_to(out[None, :, None] & out[None, None, :], (6, self.P, self.P))

    def ig_ext(self, jab_ext=None):
        """mesh%matrixIG_ext (io_mod.F90:223-250): G^-1 on the (2n+1)^2
        extended lattice of each panel, (2,2,2n+1,2n+1,6) column-major, from
        this library's conformal map or from a mesh file's jab
        (read_mesh_file_ext)."""
        m = 2 * self.n + 1
        out = np.zeros(4 * m * m * 6)
        src = None if jab_ext is None else np.ascontiguousarray(jab_ext, dtype=np.float64)
        _check("ccs_mesh_ext_metric", lib.ccs_mesh_ext_metric(self.n, self.radius, dptr(src), dptr(out)))
        return out


@dataclass
class State:
    """The prognostic state and the companions stat_field carries
    (stat_mod.F90:7-17), padded."""
    phiG: np.ndarray
    u: np.ndarray
    v: np.ndarray
    phi: np.ndarray
    uc: np.ndarray
    vc: np.ndarray
    zonal: np.ndarray | None = None
    merid: np.ndarray | None = None
    phis: np.ndarray | None = None
    extra: dict = field(default_factory=dict)

    def copy(self) -> "State":
        c = lambda a: None if a is None else a.copy()  # noqa: E731
        return State(c(self.phiG), c(self.u), c(self.v), c(self.phi), c(self.uc), c(self.vc),
                     c(self.zonal), c(self.merid), c(self.phis))


def init_case(mesh: Mesh, case_num: int = 7, seed: int = 1, n_bumps: int = 8, mountain: bool = False,
              gravity: float = GRAVITY, u0: float = 20.0, h_mean: float = 10000.0, bump_amp: float = 100.0,
              bump_rmin: float = 5.0e5, bump_rmax: float = 1.5e6, mountain_h: float = 2000.0,
              mountain_r: float = 1.5e6) -> State:
    """initTestCase (test_case_mod.F90:13-58): Williamson cases 2, 5, 6 or the
    seeded balanced jet (case 7)."""
    c = _lib.CcsCaseParams()
    lib.ccs_case_defaults(ctypes.byref(c))
    c.case_num, c.seed, c.n_bumps, c.mountain = int(case_num), int(seed), int(n_bumps), int(bool(mountain))
    c.gravity, c.u0, c.h_mean, c.bump_amp = gravity, u0, h_mean, bump_amp
    c.bump_rmin, c.bump_rmax, c.mountain_h, c.mountain_r = bump_rmin, bump_rmax, mountain_h, mountain_r
    a = {k: np.zeros(mesh.npad) for k in ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid", "phis")}
    _check("ccs_init_case", lib.ccs_init_case(
        ctypes.byref(mesh._metrics), ctypes.byref(c), *(dptr(a[k]) for k in
                                                          ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid",
                                                           "phis"))))
    return State(**a)


def _splitmix64(x):
    """SplitMix64 finaliser, elementwise on uint64 arrays (wraps mod 2^64)."""
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    return x ^ (x >> np.uint64(31))


def perturb_state(mesh: Mesh, s: State, rel: float, seed: int = 2024) -> State:
    """Copy of s whose phiG (and phi) is multiplied by 1 + rel z1 and whose
    winds (u, v and their contravariant and spherical companions) by
    1 + rel z2, with z1, z2 standard normal numbers drawn per
    cell from a hash of the cell's (lon, lat) bits and the seed.

    A halo cell holds a copy of the interior cell at the same position (its
    lon, lat are bit-for-bit copies, ghost_mod.f90:18-82), so it gets the same
    factors as its source: phiG stays a copy, and (u, v), scaled by one factor
    per cell, stays the rotation of its source's winds.  The perturbed state
    therefore keeps its halos consistent, as an FP32 rounding of the interior
    followed by the halo fill would."""
    with np.errstate(over="ignore"):
        key = (mesh.lon.view(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ mesh.lat.view(np.uint64) ^ \
            np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        out = s.copy()
        z = []
        for stream in (1, 2, 3, 4):
            r = _splitmix64(_splitmix64(key ^ np.uint64((stream * 0x632BE59BD9B4E019) & 0xFFFFFFFFFFFFFFFF)))
            z.append(((r >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53)  # uniform in (0, 1)
    rad = np.sqrt(-2.0 * np.log(z[0]))
    z1 = rad * np.cos(2.0 * np.pi * z[1])
    rad2 = np.sqrt(-2.0 * np.log(z[2]))
    z2 = rad2 * np.cos(2.0 * np.pi * z[3])
    # the companions follow: phi = phiG / sqrtG scales with phiG, and the
    # contravariant and spherical winds are linear in (u, v)
    f1, f2 = 1.0 + rel * z1, 1.0 + rel * z2
    out.phiG, out.phi = s.phiG * f1, s.phi * f1
    for k in ("u", "v", "uc", "vc", "zonal", "merid"):
        a = getattr(s, k)
        if a is not None:
            setattr(out, k, a * f2)
    return out


def cfl_time_step(mesh: Mesh, state: State, cfl: float, total_seconds: float) -> tuple[float, int]:
    """Largest dt <= cfl * min(dx / max lambda_x, dy / max lambda_y) that divides
    total_seconds, with the face eigenvalues of the Rusanov flux
    (spatial_operators_mod.F90:109-116, 134-141) on the given state."""
    n, h = mesh.n, mesh.halo
    V = mesh.view
    phit = V(state.phi + state.phis)
    ig = mesh.IG.reshape(6, mesh.P, mesh.P, 2, 2)  # [p, j, i, c, r] of column-major (r, c)
    ig11 = ig[..., 0, 0]
    ig22 = ig[..., 1, 1]
    uc, vc = V(state.uc), V(state.vc)
    rows = slice(h, h + n)
    il, ir = slice(h - 1, h + n), slice(h, h + n + 1)
    uB = 0.5 * (uc[:, rows, ir] + uc[:, rows, il])
    pB = 0.5 * (phit[:, rows, ir] + phit[:, rows, il])
    gB = 0.5 * (ig11[:, rows, ir] + ig11[:, rows, il])
    lx = np.abs(uB) + np.sqrt(gB * pB)
    vB = 0.5 * (vc[:, ir, rows] + vc[:, il, rows])
    pB = 0.5 * (phit[:, ir, rows] + phit[:, il, rows])
    gB = 0.5 * (ig22[:, ir, rows] + ig22[:, il, rows])
    ly = np.abs(vB) + np.sqrt(gB * pB)
    dt_max = cfl * min(mesh.dx / lx.max(), mesh.dx / ly.max())
    nsteps = int(math.ceil(total_seconds / dt_max - 1e-12))
    return total_seconds / nsteps, nsteps


def error_norms(x, r, area):
    """Area-weighted normalised L1, L2, Linf exactly as doc/orderCheck.m:55-74."""
    x, r, area = (np.asarray(a, dtype=np.float64).ravel() for a in (x, r, area))
    d = x - r
    l1 = np.sum(np.abs(d) * area) / np.sum(np.abs(r) * area)
    l2 = math.sqrt(np.sum(d * d * area) / np.sum(r * r * area))
    linf = np.max(np.abs(d) * area) / np.max(np.abs(r) * area)
    return float(l1), float(l2), float(linf)


class DynamicalCore:
    """One run of the GPU time-step path (a libccs handle).

    The drop-in for the reference's RK3_TVD / RK4 step (time_scheme_mod.F90)
    and spatial_operator (spatial_operators_mod.F90): all six panels stay
    resident in HBM between calls.
    """

    def __init__(self, mesh: Mesh, dt: float, precision: str = "fp64", scheme: str = "RK3_TVD",
                 device: int = 0, use_graph: bool = True, gravity: float = GRAVITY, spatial_op: str = "weno",
                 ig_ext=None):
        self.mesh = mesh
        cfg = _lib.CcsConfig()
        cfg.n, cfg.halo = mesh.n, mesh.halo
        cfg.precision = PRECISIONS[precision]
        cfg.scheme = SCHEMES[scheme]
        cfg.spatial_op = SPATIAL_OPS[spatial_op]
        cfg.eflux, cfg.device, cfg.use_graph = 0, int(device), int(bool(use_graph))
        cfg.dt, cfg.dx, cfg.radius, cfg.gravity, cfg.omega = float(dt), mesh.dx, mesh.radius, gravity, mesh.omega
        self.cfg = cfg
        self.precision, self.scheme, self.dt = precision, scheme, float(dt)
        h = ctypes.c_void_p()
        _check("ccs_create", lib.ccs_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._phis = None
        self.spatial_op = spatial_op
        if spatial_op == "weno_ext":
            self._ig_ext = np.ascontiguousarray(mesh.ig_ext() if ig_ext is None else ig_ext, dtype=np.float64)
            _check("ccs_set_mesh_ext", lib.ccs_set_mesh_ext(self._h, dptr(self._ig_ext)), self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib.ccs_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_mesh(self, phis):
        m = self.mesh
        self._phis = np.ascontiguousarray(phis, dtype=np.float64)
        _check("ccs_set_mesh", lib.ccs_set_mesh(self._h, dptr(m.sqrtG), dptr(m.G), dptr(m.IG), dptr(m.A),
                                                dptr(m.IA), dptr(m.f), dptr(self._phis)), self._h)

    def set_state(self, s: State, with_companions: bool = True):
        if self._phis is None:
            self.set_mesh(s.phis)
        args = [s.phiG, s.u, s.v] + ([s.phi, s.uc, s.vc] if with_companions else [None, None, None])
        args = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in args]
        _check("ccs_set_state", lib.ccs_set_state(self._h, *(dptr(a) for a in args)), self._h)

    def step(self, nsteps: int = 1):
        _check("ccs_step", lib.ccs_step(self._h, int(nsteps)), self._h)

    # reference-named aliases (time_scheme_mod.F90)
    def RK3_TVD(self):  # noqa: N802
        assert self.scheme == "RK3_TVD"
        self.step(1)

    def RK4(self):  # noqa: N802
        assert self.scheme == "RK4"
        self.step(1)

    def synchronize(self):
        _check("ccs_synchronize", lib.ccs_synchronize(self._h), self._h)

    def get_state(self) -> State:
        a = {k: np.zeros(self.mesh.npad) for k in ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid")}
        _check("ccs_get_state", lib.ccs_get_state(self._h, *(dptr(a[k]) for k in
                                                              ("phiG", "u", "v", "phi", "uc", "vc", "zonal",
                                                               "merid"))), self._h)
        return State(phis=self._phis, **a)

    def spatial_operator(self):
        """Tendencies (phiG, u, v) of the current state on interior cells, [panel, j, i]."""
        n = self.mesh.n
        t = [np.zeros(n * n * 6) for _ in range(3)]
        _check("ccs_tendency", lib.ccs_tendency(self._h, *(dptr(x) for x in t)), self._h)
        return tuple(x.reshape(6, n, n) for x in t)

    def diagnostics(self):
        mass, energy = ctypes.c_double(), ctypes.c_double()
        _check("ccs_diagnostics", lib.ccs_diagnostics(self._h, ctypes.byref(mass), ctypes.byref(energy)),
               self._h)
        return mass.value, energy.value

    def time_steps(self, nsteps: int, kernel_timing: bool = False):
        total, kern = ctypes.c_double(), ctypes.c_double()
        _check("ccs_time_steps", lib.ccs_time_steps(self._h, int(nsteps), ctypes.byref(total),
                                                    ctypes.byref(kern) if kernel_timing else None), self._h)
        return total.value, (kern.value if kernel_timing else None)


class Ensemble:
    """Many independent runs advanced together (the ccs_ensemble_* entry
    points): one kernel launch per RK stage covers every member's tiles.
    Members have their own mesh size, initial state and time step; each
    evolves exactly as a DynamicalCore with the same inputs would.

    tile_y: 8, 16, 32, or 0 to let the autotuner choose; persistent: 0, 1, or
    -1 to let the autotuner choose.
    """

    def __init__(self, meshes, dts, precision: str = "fp32", scheme: str = "RK3_TVD", device: int = 0,
                 use_graph: bool = True, tile_y: int = 0, persistent: int = -1, gravity: float = GRAVITY):
        meshes = list(meshes)
        if not meshes or len(meshes) != len(dts):
            raise ValueError("one time step per member")
        halo = meshes[0].halo
        if any(m.halo != halo or m.radius != meshes[0].radius or m.omega != meshes[0].omega for m in meshes):
            raise ValueError("members share halo, radius and omega")
        cfg = _lib.CcsEnsembleConfig()
        cfg.n_members, cfg.halo = len(meshes), halo
        cfg.precision, cfg.scheme = PRECISIONS[precision], SCHEMES[scheme]
        cfg.device, cfg.use_graph, cfg.tile_y, cfg.persistent = int(device), int(bool(use_graph)), int(tile_y), \
            int(persistent)
        cfg.radius, cfg.gravity, cfg.omega = meshes[0].radius, gravity, meshes[0].omega
        n = (ctypes.c_int * len(meshes))(*[m.n for m in meshes])
        dt = np.ascontiguousarray(dts, dtype=np.float64)
        dx = np.array([m.dx for m in meshes], dtype=np.float64)
        h = ctypes.c_void_p()
        _check("ccs_ensemble_create", lib.ccs_ensemble_create(ctypes.byref(cfg), n, dptr(dt), dptr(dx),
                                                              ctypes.byref(h)))
        self._h = h
        self.meshes, self.dts, self.precision, self.scheme = meshes, dt, precision, scheme
        self._phis = [None] * len(meshes)

    def _chk(self, where, code):
        if code != _lib.CCS_OK:
            raise CcsError(where, code, lib.ccs_ensemble_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib.ccs_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self.meshes)

    def set_state(self, member: int, s: State):
        m = self.meshes[member]
        self._phis[member] = np.ascontiguousarray(s.phis, dtype=np.float64)
        self._chk("ccs_ensemble_set_mesh", lib.ccs_ensemble_set_mesh(
            self._h, member, dptr(m.sqrtG), dptr(m.G), dptr(m.IG), dptr(m.A), dptr(m.IA), dptr(m.f),
            dptr(self._phis[member])))
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (s.phiG, s.u, s.v)]
        self._chk("ccs_ensemble_set_state", lib.ccs_ensemble_set_state(self._h, member, *(dptr(a) for a in args)))

    def autotune(self):
        ty, ps, ms, cached = ctypes.c_int(), ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        self._chk("ccs_ensemble_autotune", lib.ccs_ensemble_autotune(self._h, ctypes.byref(ty), ctypes.byref(ps),
                                                                     ctypes.byref(ms)))
        self._chk("ccs_ensemble_autotune_source", lib.ccs_ensemble_autotune_source(self._h, ctypes.byref(cached)))
        return {"tile_y": ty.value, "persistent": ps.value, "stage_ms": ms.value, "from_cache": bool(cached.value)}

    def step(self, nsteps: int = 1):
        """Advance every member nsteps steps of its own dt."""
        self._chk("ccs_ensemble_step", lib.ccs_ensemble_step(self._h, int(nsteps)))

    def step_members(self, nsteps):
        """Advance member k by nsteps[k] steps of its own dt, in the same
        batched launches; members that have taken their steps retire."""
        counts = [int(x) for x in nsteps]
        if len(counts) != len(self):
            raise ValueError("one step count per member")
        arr = (ctypes.c_int * len(counts))(*counts)
        self._chk("ccs_ensemble_step_members", lib.ccs_ensemble_step_members(self._h, arr))

    def run_until(self, run_time: float):
        """Advance every member to run_time seconds after its set_state:
        member k reaches int(run_time / dt[k]) steps."""
        self._chk("ccs_ensemble_run_until", lib.ccs_ensemble_run_until(self._h, float(run_time)))

    def steps_done(self, member: int) -> int:
        out = ctypes.c_longlong()
        self._chk("ccs_ensemble_steps_done", lib.ccs_ensemble_steps_done(self._h, int(member), ctypes.byref(out)))
        return out.value

    def synchronize(self):
        self._chk("ccs_ensemble_synchronize", lib.ccs_ensemble_synchronize(self._h))

    def time_steps(self, nsteps: int) -> float:
        total = ctypes.c_double()
        self._chk("ccs_ensemble_time_steps", lib.ccs_ensemble_time_steps(self._h, int(nsteps), ctypes.byref(total)))
        return total.value

    def get_state(self, member: int) -> State:
        a = {k: np.zeros(self.meshes[member].npad) for k in ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid")}
        self._chk("ccs_ensemble_get_state", lib.ccs_ensemble_get_state(
            self._h, member, *(dptr(a[k]) for k in ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid"))))
        return State(phis=self._phis[member], **a)

    def diagnostics(self, member: int):
        mass, energy = ctypes.c_double(), ctypes.c_double()
        self._chk("ccs_ensemble_diagnostics", lib.ccs_ensemble_diagnostics(self._h, member, ctypes.byref(mass),
                                                                           ctypes.byref(energy)))
        return mass.value, energy.value


def write_mesh_file(path: str, n: int):
    """Mesh file in the reference generator's schema (output_mod.f90:25-83),
    readable by the dycore's read_mesh (io_mod.F90:37-117)."""
    _check("ccs_mesh_file_write", lib.ccs_mesh_file_write(os.fsencode(path), int(n)))


def read_mesh_file(path: str):
    """(n, lon_deg, lat_deg, jab) of the cell centres of a mesh file
    (classic or 64-bit-offset netCDF), laid out as ccs_mesh_geometry's."""
    n = ctypes.c_int()
    _check("ccs_mesh_file_read", lib.ccs_mesh_file_read(os.fsencode(path), ctypes.byref(n), None, None, None))
    ni = 6 * n.value * n.value
    lon, lat, jab = np.zeros(ni), np.zeros(ni), np.zeros(6 * ni)
    _check("ccs_mesh_file_read", lib.ccs_mesh_file_read(os.fsencode(path), ctypes.byref(n), dptr(lon), dptr(lat),
                                                        dptr(jab)))
    return n.value, lon, lat, jab


def read_mesh_file_ext(path: str):
    """(n, jab_ext): the whole jab variable of a mesh file, (3,2,2n+1,2n+1,6)
    per radian, as io_mod.F90:107 reads it into mesh%jab_ext."""
    n = ctypes.c_int()
    _check("ccs_mesh_file_read_ext", lib.ccs_mesh_file_read_ext(os.fsencode(path), ctypes.byref(n), None))
    m = 2 * n.value + 1
    jab = np.zeros(36 * m * m)
    _check("ccs_mesh_file_read_ext", lib.ccs_mesh_file_read_ext(os.fsencode(path), ctypes.byref(n), dptr(jab)))
    return n.value, jab


class History:
    """History file ccsm_output.nc (io_mod.F90:309-541): history_init at
    construction, history_write_stat per output slot (1-based)."""

    def __init__(self, path: str, mesh: Mesh, phis, dt: float, case_num: int):
        d = _lib.CcsHistoryDesc(n=mesh.n, halo=mesh.halo, case_num=int(case_num), pad_=0,
                                dx_deg=90.0 / mesh.n, dt=float(dt))
        self._phis = np.ascontiguousarray(phis, dtype=np.float64)
        h = ctypes.c_void_p()
        _check("ccs_history_create", lib.ccs_history_create(os.fsencode(path), ctypes.byref(d), dptr(mesh.lon),
                                                            dptr(mesh.lat), dptr(mesh.areaCell), dptr(self._phis),
                                                            ctypes.byref(h)))
        self._h = h
        self.mesh = mesh

    def write(self, slot: int, s: State):
        arrs = [np.ascontiguousarray(getattr(s, k), dtype=np.float64) for k in
                ("phi", "u", "v", "uc", "vc", "zonal", "merid")]
        _check("ccs_history_write", lib.ccs_history_write(self._h, int(slot), dptr(arrs[0]), dptr(self._phis),
                                                          *(dptr(a) for a in arrs[1:])))

    def close(self):
        if getattr(self, "_h", None):
            _check("ccs_history_close", lib.ccs_history_close(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
