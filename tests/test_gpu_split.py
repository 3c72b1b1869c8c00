"""Split-precision path (FP32 state, FP64-derived metrics) and its FP64
instantiation, against the FP64 oracle.

The split kernels evaluate the reference equations in a non-dimensional
form with the conformal-metric identities G = sqrtG I and a shared
smoothness-indicator evaluation, so they agree with the reference to
round-off of their own precision, not bit for bit:
  - FP64 instantiation: relative differences at the 1e-12 level;
  - FP32: the reference method itself amplifies any perturbation of size
    ~1e-7 (its WENO-Z weight rule switches discontinuously,
    spatial_operators_mod.F90:348-355, most of all near the panel corners),
    so the FP32 state is held to the method's own floor: its distance from
    the FP64 oracle must stay within a small factor of the distance between
    two FP64 oracle runs whose initial states differ by FP32 rounding.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 9.80616


def _setup(ccs, n=32, halo=3, seed=3, mountain=True, hours=24.0):
    mesh = ccs.Mesh(n, halo=halo, radius=6.371e6)
    s0 = ccs.init_case(mesh, case_num=7, seed=seed, mountain=mountain)
    dt, nsteps = ccs.cfl_time_step(mesh, s0, 0.5, hours * 3600.0)
    return mesh, s0, dt, nsteps


def _nl2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


# FP32: the wind tendency of a near-balanced jet is the small residual of a
# pressure-gradient and a Coriolis term each ~1e4 times larger, so FP32
# rounding of those terms (6e-8) appears as ~1e-3 of the residual.
@pytest.mark.parametrize("precision,tol", [("fp64_fast", 1e-11), ("fp32", 2e-3)])
def test_split_tendency_close_to_oracle(ccs, orc, precision, tol):
    mesh, s0, dt, _ = _setup(ccs)
    print(f"split tendency {precision}", flush=True)
    with ccs.DynamicalCore(mesh, dt, precision=precision) as core:
        core.set_state(s0)
        tg = core.spatial_operator()
    to = orc.Oracle(mesh, s0, dt).tendency()
    for name, a, b in zip(("phiG", "u", "v"), tg, to):
        e = _nl2(a, b)
        print(f"  L({name}) normalised L2 diff {e:.3e}")
        assert e <= tol


def test_split_fp64_one_day_vs_oracle(ccs, orc):
    mesh, s0, dt, nsteps = _setup(ccs)
    print(f"split fp64_fast: 1 day C32, {nsteps} steps", flush=True)
    with ccs.DynamicalCore(mesh, dt, precision="fp64_fast") as core:
        core.set_state(s0)
        m0, _ = core.diagnostics()
        core.step(nsteps)
        s = core.get_state()
        m1, _ = core.diagnostics()
    o = orc.Oracle(mesh, s0, dt)
    o.steps(nsteps)
    area = mesh.interior(mesh.areaCell)
    l1, l2, linf = ccs.error_norms(mesh.interior(s.phi) / G, mesh.interior(o.a["phi"]) / G, area)
    print(f"  h: L1 {l1:.3e} L2 {l2:.3e} Linf {linf:.3e}; relative mass drift {(m1 - m0) / m0:.3e}")
    assert l2 <= 1e-12
    assert abs(m1 - m0) <= 1e-13 * abs(m0)


@pytest.mark.parametrize("n,hours", [(32, 24.0), (96, 6.0)])
def test_split_fp32_at_method_sensitivity_floor(ccs, orc, n, hours):
    mesh, s0, dt, nsteps = _setup(ccs, n=n, hours=hours)
    print(f"split fp32: C{n}, {hours} h, {nsteps} steps", flush=True)
    with ccs.DynamicalCore(mesh, dt, precision="fp32") as core:
        core.set_state(s0)
        m0, _ = core.diagnostics()
        core.step(nsteps)
        s = core.get_state()
        m1, _ = core.diagnostics()
    o = orc.Oracle(mesh, s0, dt)
    o.steps(nsteps)
    # floor: oracle from an initial state perturbed at FP32 rounding level
    rng = np.random.default_rng(2024)
    sp = s0.copy()
    for k in ("phiG", "u", "v"):
        a = getattr(sp, k)
        a *= 1.0 + 2.0 ** -24 * rng.standard_normal(a.shape)
    op = orc.Oracle(mesh, sp, dt)
    op.steps(nsteps)
    area = mesh.interior(mesh.areaCell)
    ref = mesh.interior(o.a["phi"]) / G
    e32 = ccs.error_norms(mesh.interior(s.phi) / G, ref, area)
    efl = ccs.error_norms(mesh.interior(op.a["phi"]) / G, ref, area)
    print(f"  fp32 vs oracle      h: L1 {e32[0]:.3e} L2 {e32[1]:.3e} Linf {e32[2]:.3e}")
    print(f"  perturbed vs oracle h: L1 {efl[0]:.3e} L2 {efl[1]:.3e} Linf {efl[2]:.3e}")
    print(f"  relative mass drift {(m1 - m0) / m0:.3e}")
    assert e32[1] <= 4.0 * efl[1] + 1e-9
    # mass: exact flux telescoping, FP64 compensated sums and the rebase of
    # the FP32 deviation every 16 steps leave the round-off of the increments
    # (measured: C32 over 24 h 2.3e-10, C96 over 6 h 3.2e-11)
    assert abs(m1 - m0) <= 5e-10 * abs(m0)


def orc_zonal(mesh, o):
    A = mesh.A.reshape(-1, 2, 2)
    IG = mesh.IG.reshape(-1, 2, 2)
    u, v = o.a["u"], o.a["v"]
    # covProjPlane2Sphere with matrix = matmul(A, IG); arrays are [cell, c, r]
    M11 = A[:, 0, 0] * IG[:, 0, 0] + A[:, 1, 0] * IG[:, 0, 1]
    M12 = A[:, 0, 0] * IG[:, 1, 0] + A[:, 1, 0] * IG[:, 1, 1]
    return M11 * u + M12 * v


def test_split_fp64_matches_parity_path_after_steps(ccs):
    mesh, s0, dt, _ = _setup(ccs, n=48, halo=4, seed=9)
    res = {}
    for prec in ("fp64", "fp64_fast"):
        print(f"{prec}: 20 steps C48 halo 4", flush=True)
        with ccs.DynamicalCore(mesh, dt, precision=prec) as core:
            core.set_state(s0)
            core.step(20)
            res[prec] = core.get_state()
    mask = ~mesh.corner_mask().ravel()
    for k in ("phiG", "u", "v"):
        a, b = getattr(res["fp64_fast"], k)[mask], getattr(res["fp64"], k)[mask]
        e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        print(f"  {k}: max rel diff {e:.3e}")
        assert e <= 1e-11


def test_split_rk4_runs_and_conserves_mass(ccs):
    mesh, s0, dt, _ = _setup(ccs, n=32, seed=11)
    with ccs.DynamicalCore(mesh, dt, precision="fp32", scheme="RK4") as core:
        core.set_state(s0)
        m0, _ = core.diagnostics()
        core.step(50)
        m1, _ = core.diagnostics()
    assert abs(m1 - m0) <= 1e-9 * abs(m0)


def test_split_rejects_nonconformal_grid(ccs):
    mesh, s0, dt, _ = _setup(ccs, n=16)
    bad = mesh.IG.copy()
    bad[1::4] += 1e-3 * np.abs(bad[0::4])  # IG21 off-diagonal: not conformal
    with ccs.DynamicalCore(mesh, dt, precision="fp32") as core:
        mesh_ig = mesh.IG
        try:
            mesh.IG = bad
            with pytest.raises(ccs.CcsError):
                core.set_mesh(s0.phis)
        finally:
            mesh.IG = mesh_ig


@pytest.mark.parametrize("scheme", ["RK3_TVD", "RK4"])
def test_split_step_grouping_does_not_change_results(ccs, scheme):
    # the geopotential rebase runs after every 16th step counted from
    # set_state, whatever the grouping of steps into calls (16-step graphs,
    # single-step graphs, a trailing rebase graph), so 45 steps give the same
    # state bit for bit in one call, in five uneven calls and in 45 calls
    mesh, s0, dt, _ = _setup(ccs, n=40, seed=6)
    out = []
    for groups in ((45,), (1, 2, 15, 17, 10), (1,) * 45):
        print(f"fp32 {scheme}: 45 steps as {len(groups)} calls", flush=True)
        with ccs.DynamicalCore(mesh, dt, precision="fp32", scheme=scheme) as core:
            core.set_state(s0)
            for g in groups:
                core.step(g)
            out.append((core.get_state(), core.diagnostics()))
    for s, d in out[1:]:
        for k in ("phiG", "u", "v"):
            assert np.array_equal(getattr(s, k), getattr(out[0][0], k)), k
        assert d == out[0][1]


@pytest.mark.parametrize("precision,scheme,n", [("fp32", "RK3_TVD", 40), ("fp32", "RK3_TVD", 100),
                                                ("fp32", "RK4", 40), ("fp64_fast", "RK3_TVD", 40)])
def test_split_results_do_not_depend_on_the_tile_height(ccs, monkeypatch, precision, scheme, n):
    # a face value is computed by the tile on either side of it, by the
    # interior windows of one and the boundary windows of the other; both
    # round alike, so 8-, 16- and 32-row tiles give the same state bit for
    # bit (n = 100: partial tiles in both directions)
    mesh, s0, dt, _ = _setup(ccs, n=n, seed=8)
    out = []
    for ty in (16, 8, 32):
        monkeypatch.setenv("CCS_SPLIT_TILE_Y", str(ty))
        print(f"{precision} {scheme} C{n}: 21 steps with {ty}-row tiles", flush=True)
        with ccs.DynamicalCore(mesh, dt, precision=precision, scheme=scheme) as core:
            core.set_state(s0)
            core.step(21)
            out.append(core.get_state())
    monkeypatch.delenv("CCS_SPLIT_TILE_Y")
    for s in out[1:]:
        for k in ("phiG", "u", "v"):
            assert np.array_equal(getattr(s, k), getattr(out[0], k)), k


def test_split_halo_reference_follows_its_source_cell(ccs):
    # An uploaded state whose halo phiG differs from the interior cell it
    # copies (here by 1e-3 relative).  Both paths read the halo as uploaded in
    # the first stage and hold copies of the source cells afterwards
    # (fill_ghost, ghost_mod.f90:163-170): the split path keeps the halo's
    # FP64 reference value equal to its source's, so FP64-fast and the FP64
    # parity path stay together to round-off.
    mesh, s0, dt, _ = _setup(ccs, n=32, seed=5)
    sp = s0.copy()
    halo = np.zeros((6, mesh.P, mesh.P), bool)
    h, n = mesh.halo, mesh.n
    halo[:, :, :] = True
    halo[:, h:h + n, h:h + n] = False
    halo &= ~mesh.corner_mask()
    sp.phiG[halo.ravel()] *= 1.0 + 1e-3
    res = {}
    for prec in ("fp64", "fp64_fast"):
        with ccs.DynamicalCore(mesh, dt, precision=prec) as core:
            core.set_state(sp, with_companions=False)
            core.step(5)
            res[prec] = core.get_state()
    mask = ~mesh.corner_mask().ravel()
    for k in ("phiG", "u", "v"):
        a, b = getattr(res["fp64_fast"], k)[mask], getattr(res["fp64"], k)[mask]
        e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        print(f"  {k}: max rel diff {e:.3e}", flush=True)
        assert e <= 1e-11


@pytest.mark.parametrize("n,hours", [(48, 24.0), (96, 12.0)])
def test_case2_steady_flow_fp32_vs_fp64(ccs, n, hours):
    # Williamson case 2 (test_case_mod.F90:61-95): a steady zonal flow, so the
    # rounding of the FP32 path is not amplified by developing structure.
    mesh = ccs.Mesh(n, halo=3, radius=6.371e6)
    s0 = ccs.init_case(mesh, case_num=2)
    dt, nsteps = ccs.cfl_time_step(mesh, s0, 0.5, hours * 3600.0)
    out = {}
    for name, prec, st0 in (("fp32", "fp32", s0), ("fp64", "fp64", s0),
                            ("floor", "fp64", ccs.perturb_state(mesh, s0, 2.0 ** -24))):
        with ccs.DynamicalCore(mesh, dt, precision=prec) as core:
            core.set_state(st0)
            m0, _ = core.diagnostics()
            core.step(nsteps)
            m1, _ = core.diagnostics()
            out[name] = (core.get_state(), (m1 - m0) / m0)
    area = mesh.interior(mesh.areaCell)
    ref = mesh.interior(out["fp64"][0].phi) / G
    e32 = ccs.error_norms(mesh.interior(out["fp32"][0].phi) / G, ref, area)
    efl = ccs.error_norms(mesh.interior(out["floor"][0].phi) / G, ref, area)
    print(f"case 2 C{n}, {hours} h ({nsteps} steps): L2(h) fp32 {e32[1]:.3e} (floor {efl[1]:.3e}), "
          f"Linf {e32[2]:.3e}; mass drift fp32 {out['fp32'][1]:.3e}", flush=True)
    assert e32[1] <= 4.0 * efl[1] + 1e-9
    assert abs(out["fp32"][1]) <= 1e-10


def test_handle_created_while_another_run_is_in_flight(ccs):
    # A handle is created, cleared and uploaded while a long run of another
    # handle still occupies the GPU (its steps are asynchronous): the new
    # handle's buffers must be cleared before its uploads land, whatever the
    # other streams are doing, so its run equals the same run on an idle GPU.
    mesh, s0, dt, _ = _setup(ccs, n=24, seed=21)
    big, sb, dtb, _ = _setup(ccs, n=360, seed=22, hours=6.0)
    with ccs.DynamicalCore(mesh, dt, precision="fp32") as core:
        core.set_state(s0)
        core.step(30)
        idle = core.get_state()
    with ccs.DynamicalCore(big, dtb, precision="fp32") as busy:
        busy.set_state(sb)
        busy.step(400)  # not synchronised: still running below
        for prec in ("fp32", "fp64"):
            with ccs.DynamicalCore(mesh, dt, precision=prec) as core:
                core.set_state(s0)
                core.step(30)
                s = core.get_state()
                assert np.all(np.isfinite(mesh.interior(s.phi)))
                if prec == "fp32":
                    for k in ("phiG", "u", "v"):
                        assert np.array_equal(getattr(s, k), getattr(idle, k)), k
        busy.synchronize()
