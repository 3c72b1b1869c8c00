"""Golden fixtures written from the reference's own compiled code
(tests/golden/make_golden.py, golden_c6.npz): C6, halo 3, Williamson case 6.

- CPU: the oracle, fed the fixture's inputs, reproduces the reference's
  spatial_operator and 3 SSP-RK3 steps bit for bit for every
  spatial_operators_mod source; this library's conformal map reproduces the
  generator's xmtoxc on the extended lattice to a few ulp.
- GPU: the product (FP64 parity path, through the C ABI with the fixture's
  tables) reproduces the same reference outputs bit for bit.
These tests need neither the reference tree nor oracle/_ref.
"""
import ctypes
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "golden_c6.npz")
OPS = ["weno", "centred2", "centred4", "centred6", "weno_ext"]
STATE = ("phiG", "u", "v", "phi", "uc", "vc")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


class _Mesh:
    """The fixture's tables in the shape oracle_lib expects of a Mesh."""

    def __init__(self, g):
        self.n, self.halo, self.dx = int(g["n"]), int(g["halo"]), float(g["dx"])
        self.P = self.n + 2 * self.halo
        self.npad = self.P * self.P * 6
        for k in ("sqrtG", "G", "IG", "A", "IA", "f"):
            setattr(self, k, np.ascontiguousarray(g["in_" + k]))

    def interior(self, a):
        h, n = self.halo, self.n
        return a.reshape(6, self.P, self.P)[:, h:h + n, h:h + n]

    def noncorner(self):
        idx = np.arange(self.P) - self.halo
        out = (idx < 0) | (idx >= self.n)
        return ~(out[None, :, None] & out[None, None, :]).repeat(6, 0).ravel()


class _State:
    def __init__(self, g):
        for k in STATE + ("phis",):
            setattr(self, k, np.ascontiguousarray(g["in_" + k]))
        self.zonal = self.merid = None


@pytest.mark.parametrize("op", OPS)
def test_oracle_reproduces_reference_golden(orc, gold, op):
    m, s = _Mesh(gold), _State(gold)
    o = orc.Oracle(m, s, float(gold["dt"]), op=op, ig_ext=gold["in_ig_ext"])
    for k, t in zip(("phiG", "u", "v"), o.tendency()):
        assert np.array_equal(t, gold[f"{op}_tend_{k}"]), k
    o.steps(3)
    mask = m.noncorner()
    for k in STATE:
        assert np.array_equal(o.a[k][mask], gold[f"{op}_rk3_{k}"][mask]), k


def test_map_reproduces_reference_generator_golden(ccs, gold, tmp_path):
    from scipy.io import netcdf_file

    path = str(tmp_path / "mesh_C6.nc")
    ccs.write_mesh_file(path, 6)
    with netcdf_file(path, "r", mmap=False) as f:
        X = np.stack([f.variables[k][:].astype(np.float64) for k in ("x", "y", "z")], -1)
        jab = f.variables["jab"][:].astype(np.float64)
    assert np.max(np.abs(X - gold["map_xc"])) <= 2e-15
    ref_jab = gold["map_dxcdxm"] / 45.0 / (np.pi / 180.0)  # mesh_mod.f90:163
    assert np.max(np.abs(jab - ref_jab)) <= 5e-14 * np.max(np.abs(ref_jab))


@pytest.mark.gpu
@pytest.mark.parametrize("op", OPS)
def test_gpu_reproduces_reference_golden(ccs, gold, op):
    from conformal_cubed_sphere_amd import _lib

    L, d = _lib.lib, _lib.dptr
    m = _Mesh(gold)
    cfg = _lib.CcsConfig(n=m.n, halo=m.halo, precision=_lib.CCS_FP64_PARITY, scheme=_lib.CCS_RK3_TVD, eflux=0,
                         device=0, use_graph=1, spatial_op=ccs.SPATIAL_OPS[op], dt=float(gold["dt"]), dx=m.dx,
                         radius=6371229.0, gravity=9.80616, omega=7.292e-5)
    h = ctypes.c_void_p()
    print(f"golden C6 op={op}: tendency and 3 RK3 steps", flush=True)
    assert L.ccs_create(ctypes.byref(cfg), ctypes.byref(h)) == 0
    try:
        ins = {k: np.ascontiguousarray(gold["in_" + k]) for k in
               ("sqrtG", "G", "IG", "A", "IA", "f", "phis", "ig_ext") + STATE}
        assert L.ccs_set_mesh(h, *(d(ins[k]) for k in ("sqrtG", "G", "IG", "A", "IA", "f", "phis"))) == 0
        if op == "weno_ext":
            assert L.ccs_set_mesh_ext(h, d(ins["ig_ext"])) == 0
        assert L.ccs_set_state(h, *(d(ins[k]) for k in STATE)) == 0
        ni = m.n * m.n * 6
        t = [np.zeros(ni) for _ in range(3)]
        assert L.ccs_tendency(h, *(d(x) for x in t)) == 0
        for k, x in zip(("phiG", "u", "v"), t):
            assert np.array_equal(x.reshape(6, m.n, m.n), gold[f"{op}_tend_{k}"]), k
        assert L.ccs_step(h, 3) == 0
        out = {k: np.zeros(m.npad) for k in STATE}
        assert L.ccs_get_state(h, *(d(out[k]) for k in STATE), None, None) == 0
        mask = m.noncorner()
        for k in STATE:
            assert np.array_equal(out[k][mask], gold[f"{op}_rk3_{k}"][mask]), k
    finally:
        L.ccs_destroy(h)
