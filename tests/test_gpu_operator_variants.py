"""FP64 parity of the spatial-operator variants on the GPU against the oracle.

The reference can be built with spatial_operators_mod_{2nd,4th,6th,WENO}.F90
in place of spatial_operators_mod.F90 (ccs_config.spatial_op selects the
matching kernel).  The oracle's variants equal the reference's own compiled
modules bit for bit (tests/test_oracle_vs_reference.py); here the GPU kernels
must equal the oracle bit for bit: the tendency, RK3 and RK4 steps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPS = ["centred2", "centred4", "centred6", "weno_ext"]


def _setup(ccs, n, halo, case_num=7, mountain=True, seed=5):
    mesh = ccs.Mesh(n, halo=halo)
    s0 = ccs.init_case(mesh, case_num=case_num, seed=seed, mountain=mountain)
    dt, _ = ccs.cfl_time_step(mesh, s0, 0.5, 3600.0)
    return mesh, s0, dt


@pytest.mark.parametrize("halo", [3, 4])
@pytest.mark.parametrize("op", OPS)
def test_variant_tendency_bitwise(ccs, orc, op, halo):
    mesh, s0, dt = _setup(ccs, 32, halo)
    print(f"tendency C32 halo={halo} op={op}", flush=True)
    with ccs.DynamicalCore(mesh, dt, spatial_op=op) as core:
        core.set_state(s0)
        tg = core.spatial_operator()
    to = orc.Oracle(mesh, s0, dt, op=op).tendency()
    for name, a, b in zip(("phiG", "u", "v"), tg, to):
        print(f"  L({name}): bitwise-equal fraction {np.mean(a == b):.6f}, "
              f"max rel diff {np.max(np.abs(a - b)) / np.max(np.abs(b)):.3e}")
        assert np.array_equal(a, b)


@pytest.mark.parametrize("op,scheme,nsteps", [(op, "RK3_TVD", 12) for op in OPS] + [("weno_ext", "RK4", 5),
                                                                                     ("centred4", "RK4", 5)])
def test_variant_steps_bitwise(ccs, orc, op, scheme, nsteps):
    mesh, s0, dt = _setup(ccs, 32, 3, case_num=7, mountain=True, seed=2)
    mask = ~mesh.corner_mask().ravel()
    print(f"{scheme} C32 op={op}: {nsteps} steps", flush=True)
    with ccs.DynamicalCore(mesh, dt, scheme=scheme, spatial_op=op) as core:
        core.set_state(s0)
        core.step(nsteps)
        s = core.get_state()
    o = orc.Oracle(mesh, s0, dt, op=op)
    o.steps(nsteps, scheme=0 if scheme == "RK3_TVD" else 1)
    for k in ("phiG", "u", "v", "phi", "uc", "vc"):
        a, b = getattr(s, k)[mask], o.a[k][mask]
        print(f"  {k}: bitwise-equal fraction {np.mean(a == b):.6f}")
        assert np.array_equal(a, b)


def test_weno_ext_from_mesh_file_metric(ccs, orc, tmp_path):
    # the extended-lattice metric from a mesh file's jab gives the same run
    mesh, s0, dt = _setup(ccs, 16, 3, case_num=2, mountain=False)
    path = str(tmp_path / "mesh_C16.nc")
    ccs.write_mesh_file(path, 16)
    n, jab_ext = ccs.read_mesh_file_ext(path)
    ig = mesh.ig_ext(jab_ext)
    print("weno_ext C16 from mesh file: 4 steps", flush=True)
    with ccs.DynamicalCore(mesh, dt, spatial_op="weno_ext", ig_ext=ig) as core:
        core.set_state(s0)
        core.step(4)
        s = core.get_state()
    o = orc.Oracle(mesh, s0, dt, op="weno_ext", ig_ext=ig)
    o.steps(4)
    mask = ~mesh.corner_mask().ravel()
    for k in ("phiG", "u", "v"):
        assert np.array_equal(getattr(s, k)[mask], o.a[k][mask])


# ---- the centred operators on the split precisions
# CCS_FP32_SPLIT and CCS_FP64_FAST run the centred operators with the split
# path's arithmetic (geopotential as a deviation from a reference field,
# contravariant winds U/sqrtg on the conformal grid), so they are close to the
# oracle, not bitwise.  The fp64_fast bound is far below the distance between
# two different operators, so it also pins which operator ran.
CENTRED = ["centred2", "centred4", "centred6"]


def _nl2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


@pytest.mark.parametrize("precision,tol", [("fp64_fast", 1e-11), ("fp32", 2e-3)])
@pytest.mark.parametrize("op", CENTRED)
def test_split_centred_tendency_close_to_oracle(ccs, orc, op, precision, tol):
    mesh, s0, dt = _setup(ccs, 32, 3)
    print(f"split tendency C32 {precision} op={op}", flush=True)
    with ccs.DynamicalCore(mesh, dt, precision=precision, spatial_op=op) as core:
        core.set_state(s0)
        tg = core.spatial_operator()
    to = orc.Oracle(mesh, s0, dt, op=op).tendency()
    tw = orc.Oracle(mesh, s0, dt, op="weno").tendency()
    for name, a, b, w in zip(("phiG", "u", "v"), tg, to, tw):
        e, ew = _nl2(a, b), _nl2(a, w)
        print(f"  L({name}) normalised L2 diff {e:.3e} (against the WENO operator {ew:.3e})")
        assert e <= tol
        assert ew >= 1e-6


@pytest.mark.parametrize("precision,scheme,nsteps,tol", [
    ("fp64_fast", "RK3_TVD", 12, 1e-11), ("fp64_fast", "RK4", 5, 1e-11),
    ("fp32", "RK3_TVD", 12, 1e-5), ("fp32", "RK4", 5, 1e-5)])
@pytest.mark.parametrize("op", CENTRED)
def test_split_centred_steps_close_to_parity_path(ccs, op, precision, scheme, nsteps, tol):
    # against the FP64 parity path of the same operator, which equals the
    # oracle bit for bit (test_variant_steps_bitwise)
    mesh, s0, dt = _setup(ccs, 32, 3, case_num=7, mountain=True, seed=2)
    print(f"{scheme} C32 {precision} op={op}: {nsteps} steps", flush=True)
    states = {}
    for prec in (precision, "fp64"):
        with ccs.DynamicalCore(mesh, dt, precision=prec, scheme=scheme, spatial_op=op) as core:
            core.set_state(s0)
            core.step(nsteps)
            states[prec] = core.get_state()
    h = mesh.interior
    for k in ("phi", "u", "v"):
        e = _nl2(h(getattr(states[precision], k)), h(getattr(states["fp64"], k)))
        print(f"  {k}: normalised L2 diff {e:.3e}")
        assert e <= tol


@pytest.mark.parametrize("op", ["weno"] + OPS)
def test_variant_steps_bitwise_at_c96(ccs, orc, op):
    # a larger grid than the C32 cases above (3 tiles per panel row, ragged
    # last tile column), the benchmark recipe with the mountain, 3 RK3 steps
    mesh = ccs.Mesh(96, halo=3, radius=6.371e6, omega=7.292e-5)
    s0 = ccs.init_case(mesh, case_num=7, seed=1, n_bumps=8, mountain=True)
    dt, _ = ccs.cfl_time_step(mesh, s0, 0.5, 86400.0)
    print(f"RK3 C96 op={op}: 3 steps", flush=True)
    with ccs.DynamicalCore(mesh, dt, spatial_op=op) as core:
        core.set_state(s0)
        core.step(3)
        s = core.get_state()
    o = orc.Oracle(mesh, s0, dt, op=op)
    o.steps(3)
    mask = ~mesh.corner_mask().ravel()
    for k in ("phiG", "u", "v", "phi", "uc", "vc"):
        assert np.array_equal(getattr(s, k)[mask], o.a[k][mask]), k
