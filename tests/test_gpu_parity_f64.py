"""FP64 parity of the GPU time-step path against the CPU oracle.

The GPU path (libccs, CCS_FP64_PARITY) and the oracle (a restatement of the
reference's time_scheme_mod / spatial_operators_mod / ghost_mod) run on the
same seeded inputs.  Both evaluate the reference's expressions in the same
order with IEEE FP64 rounding, so the bar is round-off: relative differences
<= 1e-13 per field after a few steps and normalised L2(h) <= 1e-12 after a
simulated day (the bar of the C180 configuration).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b, mask):
    a, b = a[mask], b[mask]
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _setup(ccs, n, halo, case_num=7, mountain=False, seed=1, hours=24.0):
    mesh = ccs.Mesh(n, halo=halo)
    s0 = ccs.init_case(mesh, case_num=case_num, seed=seed, mountain=mountain)
    dt, nsteps = ccs.cfl_time_step(mesh, s0, 0.5, hours * 3600.0)
    return mesh, s0, dt, nsteps


@pytest.mark.parametrize("halo", [3, 4])
@pytest.mark.parametrize("case_num,mountain", [(7, False), (7, True), (2, False), (5, False), (6, False)])
def test_tendency_matches_oracle(ccs, orc, halo, case_num, mountain):
    mesh, s0, dt, _ = _setup(ccs, 32, halo, case_num, mountain, seed=5)
    print(f"tendency C32 halo={halo} case={case_num} mountain={mountain}", flush=True)
    with ccs.DynamicalCore(mesh, dt) as core:
        core.set_state(s0)
        tg = core.spatial_operator()
    to = orc.Oracle(mesh, s0, dt).tendency()
    for name, a, b in zip(("phiG", "u", "v"), tg, to):
        scale = np.max(np.abs(b))
        err = np.max(np.abs(a - b)) / scale
        print(f"  L({name}): max rel diff {err:.3e}, bitwise-equal fraction {np.mean(a == b):.6f}")
        assert err <= 1e-14


@pytest.mark.parametrize("halo", [3, 4])
def test_rk3_steps_match_oracle(ccs, orc, halo):
    mesh, s0, dt, _ = _setup(ccs, 32, halo, 7, True, seed=2)
    mask = ~mesh.corner_mask().ravel()
    orc_run = orc.Oracle(mesh, s0, dt)
    with ccs.DynamicalCore(mesh, dt) as core:
        core.set_state(s0)
        for nst in (1, 9):
            print(f"RK3 C32 halo={halo}: +{nst} steps", flush=True)
            core.step(nst)
            orc_run.steps(nst)
            s = core.get_state()
            for k in ("phiG", "u", "v", "phi", "uc", "vc"):
                err = _rel(getattr(s, k), orc_run.a[k], mask)
                print(f"  {k}: max rel diff {err:.3e}")
                assert err <= 1e-13


def test_rk4_steps_match_oracle(ccs, orc):
    mesh, s0, dt, _ = _setup(ccs, 32, 4, 6, False, seed=1)
    mask = ~mesh.corner_mask().ravel()
    print("RK4 C32 halo=4 case 6: 5 steps", flush=True)
    with ccs.DynamicalCore(mesh, dt, scheme="RK4") as core:
        core.set_state(s0)
        core.step(5)
        s = core.get_state()
    o = orc.Oracle(mesh, s0, dt)
    o.steps(5, scheme=1)
    for k in ("phiG", "u", "v"):
        err = _rel(getattr(s, k), o.a[k], mask)
        print(f"  {k}: max rel diff {err:.3e}")
        assert err <= 1e-13


@pytest.mark.parametrize("precision,scheme", [("fp64", "RK3_TVD"), ("fp64", "RK4"), ("fp32", "RK3_TVD")])
def test_graph_replay_equals_eager(ccs, precision, scheme):
    """One-step and 16-step graphs, replayed several times with host work and
    diagnostics between the calls, give the eager launches' state bit for bit."""
    mesh, s0, dt, _ = _setup(ccs, 48, 3, 7, True, seed=4)
    out = []
    for use_graph in (False, True):
        print(f"{precision} {scheme} graph={use_graph}: 43 steps in 5 calls", flush=True)
        with ccs.DynamicalCore(mesh, dt, precision=precision, scheme=scheme, use_graph=use_graph) as core:
            core.set_state(s0)
            diag = []
            for n in (1, 2, 3, 17, 20):
                core.step(n)
                diag.append(core.diagnostics())
            out.append((core.get_state(), diag))
    for k in ("phiG", "u", "v"):
        assert np.array_equal(getattr(out[0][0], k), getattr(out[1][0], k)), k
    assert out[0][1] == out[1][1]


def test_diagnostics_compensated(ccs, orc):
    mesh, s0, dt, _ = _setup(ccs, 32, 3, 7, True, seed=7)
    o = orc.Oracle(mesh, s0, dt)
    with ccs.DynamicalCore(mesh, dt) as core:
        core.set_state(s0)
        m0, e0 = core.diagnostics()
        _, _, om, oe = o.diagnostics()
        print(f"initial mass {m0:.17e} oracle {om:.17e}; energy {e0:.17e} oracle {oe:.17e}", flush=True)
        assert abs(m0 - om) <= 2e-16 * abs(om)
        assert abs(e0 - oe) <= 4e-16 * abs(oe)
        core.step(20)
        o.steps(20)
        m1, e1 = core.diagnostics()
        _, _, om, oe = o.diagnostics()
        assert abs(m1 - om) <= 1e-14 * abs(om)
        assert abs(e1 - oe) <= 1e-14 * abs(oe)
        # flux-form mass conservation: only round-off drift
        assert abs(m1 - m0) <= 1e-13 * abs(m0)


def test_one_day_c32_normalised_l2(ccs, orc):
    mesh, s0, dt, nsteps = _setup(ccs, 32, 3, 7, True, seed=1)
    print(f"1 day C32: {nsteps} steps of dt={dt:.3f}", flush=True)
    with ccs.DynamicalCore(mesh, dt) as core:
        core.set_state(s0)
        core.step(nsteps)
        s = core.get_state()
    o = orc.Oracle(mesh, s0, dt)
    o.steps(nsteps)
    area = mesh.interior(mesh.areaCell)
    h_gpu = mesh.interior(s.phi) / 9.80616
    h_orc = mesh.interior(o.a["phi"]) / 9.80616
    l1, l2, linf = ccs.error_norms(h_gpu, h_orc, area)
    print(f"  normalised L1/L2/Linf(h) = {l1:.3e} {l2:.3e} {linf:.3e}")
    assert l2 <= 1e-12


def test_reference_namelist_configuration_one_day(ccs, orc):
    """The configuration of the reference's shipped namelist
    (dynamic_core/run/namelist.input): RK4, Williamson case 2, halo 4 and a
    fixed dt = 100 s, here at C48 for one simulated day (864 steps), equal
    to the oracle bit for bit away from the never-filled halo corners."""
    mesh = ccs.Mesh(48, halo=4)
    s0 = ccs.init_case(mesh, case_num=2)
    dt, nsteps = 100.0, 864
    print(f"reference namelist: C48 halo 4, case 2, RK4, {nsteps} steps of dt={dt}", flush=True)
    with ccs.DynamicalCore(mesh, dt, scheme="RK4") as core:
        core.set_state(s0)
        core.step(nsteps)
        s = core.get_state()
        m, e = core.diagnostics()
    o = orc.Oracle(mesh, s0, dt)
    o.steps(nsteps, scheme=1)
    mask = ~mesh.corner_mask().ravel()
    for k in ("phiG", "u", "v", "phi"):
        same = np.array_equal(getattr(s, k)[mask], o.a[k][mask])
        print(f"  {k}: bitwise equal {same}, max rel diff {_rel(getattr(s, k), o.a[k], mask):.3e}", flush=True)
        assert same, k
    _, _, om, oe = o.diagnostics()
    assert abs(m - om) <= 1e-14 * abs(om) and abs(e - oe) <= 1e-14 * abs(oe)


@pytest.mark.slow
def test_one_day_c180_matches_cpu_oracle(ccs, orc):
    """The drop-in FP64 path at C180 (the benchmark recipe: a = 6.371e6 m,
    seeded jet + 8 bumps, flat bottom, SSP-RK3 at CFL 0.5) against the CPU
    oracle after one simulated day: normalised L2(h) <= 1e-12 (in practice
    the two are equal bit for bit)."""
    mesh = ccs.Mesh(180, halo=3, radius=6.371e6, omega=7.292e-5)
    s0 = ccs.init_case(mesh, case_num=7, seed=1, n_bumps=8, mountain=False)
    dt, nsteps = ccs.cfl_time_step(mesh, s0, 0.5, 86400.0)
    print(f"1 day C180: {nsteps} steps of dt={dt:.3f}", flush=True)
    with ccs.DynamicalCore(mesh, dt) as core:
        core.set_state(s0)
        core.step(nsteps)
        s = core.get_state()
    o = orc.Oracle(mesh, s0, dt)
    o.steps(nsteps)
    area = mesh.interior(mesh.areaCell)
    l1, l2, linf = ccs.error_norms(mesh.interior(s.phi) / 9.80616, mesh.interior(o.a["phi"]) / 9.80616, area)
    print(f"  normalised L1/L2/Linf(h) = {l1:.3e} {l2:.3e} {linf:.3e}", flush=True)
    assert l2 <= 1e-12
    mask = ~mesh.corner_mask().ravel()
    print(f"  bitwise equal: {np.array_equal(s.phiG[mask], o.a['phiG'][mask])}", flush=True)
