"""Batched ensemble runs with per-member step counts (ccs_ensemble_step_members,
ccs_ensemble_run_until), members retiring from the tile list, per-member
rebase schedules, the persistent grid's tile partition at every grid size,
the autotune cache, and members against the FP64 oracle.

The reference runs one member per process and derives its step count from
its own time step (parameters_mod.F90:149, nsteps = total_run_time / dt); a
member here must end bit for bit where that single run ends."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 9.80616


def _members(ccs, sizes, hours=6.0, seed0=11):
    meshes, states, dts = [], [], []
    cache = {}
    for k, n in enumerate(sizes):
        if n not in cache:
            cache[n] = ccs.Mesh(n, halo=3, radius=6.371e6)
        m = cache[n]
        s = ccs.init_case(m, case_num=7, seed=seed0 + k, mountain=(k % 2 == 0))
        dt, _ = ccs.cfl_time_step(m, s, 0.5, hours * 3600.0)
        meshes.append(m)
        states.append(s)
        dts.append(dt)
    return meshes, states, dts


def _single(ccs, mesh, s0, dt, nsteps, precision="fp32"):
    with ccs.DynamicalCore(mesh, dt, precision=precision, use_graph=True) as core:
        core.set_state(s0)
        core.step(nsteps)
        return core.get_state(), core.diagnostics()


def _same(a, b):
    return all(np.array_equal(getattr(a, f), getattr(b, f)) for f in ("phiG", "u", "v"))


def test_run_until_takes_each_member_to_its_own_step_count(ccs):
    sizes = (16, 24, 32, 20, 48)
    meshes, states, dts = _members(ccs, sizes)
    with ccs.Ensemble(meshes, dts, precision="fp32", tile_y=16, persistent=0) as ens:
        for k, s in enumerate(states):
            ens.set_state(k, s)
        for hours in (3.0, 7.5):
            t = hours * 3600.0
            ens.run_until(t)
            want = [int(t / dt) for dt in dts]
            print(f"run_until {hours} h: steps {want}", flush=True)
            assert len(set(want)) > 1  # the members really take different step counts
            for k in range(len(ens)):
                assert ens.steps_done(k) == want[k]
                ref, (m1, e1) = _single(ccs, meshes[k], states[k], dts[k], want[k])
                assert _same(ens.get_state(k), ref), k
                assert ens.diagnostics(k) == (m1, e1)


@pytest.mark.parametrize("persistent", [0, 1])
def test_step_members_with_uneven_counts(ccs, persistent):
    sizes = (24, 16, 32, 20)
    counts = (5, 0, 37, 18)   # 0: a member that does not step at all
    meshes, states, dts = _members(ccs, sizes, seed0=31)
    print(f"step_members {counts}, persistent={persistent}", flush=True)
    with ccs.Ensemble(meshes, dts, precision="fp32", tile_y=16, persistent=persistent) as ens:
        for k, s in enumerate(states):
            ens.set_state(k, s)
        ens.step_members(counts)
        for k in range(len(ens)):
            assert ens.steps_done(k) == counts[k]
            ref, _ = _single(ccs, meshes[k], states[k], dts[k], counts[k])
            assert _same(ens.get_state(k), ref), k


def test_reupload_of_one_member_keeps_the_others_rebase_schedule(ccs):
    # member 0 keeps stepping across the re-upload of member 1: its rebase
    # points (every 16 of its own steps) must not move, and member 1's must
    # restart from its new upload
    sizes = (24, 32)
    meshes, states, dts = _members(ccs, sizes, seed0=41)
    with ccs.Ensemble(meshes, dts, precision="fp32", tile_y=16, persistent=0) as ens:
        for k, s in enumerate(states):
            ens.set_state(k, s)
        ens.step(7)
        ens.set_state(1, states[1])
        ens.step(20)
        assert ens.steps_done(0) == 27 and ens.steps_done(1) == 20
        ref0, _ = _single(ccs, meshes[0], states[0], dts[0], 27)
        ref1, _ = _single(ccs, meshes[1], states[1], dts[1], 20)
        assert _same(ens.get_state(0), ref0)
        assert _same(ens.get_state(1), ref1)


def test_one_member_persistent_grid_smaller_than_eight_workgroups(ccs, monkeypatch):
    # n = 24 with 32-row tiles: 6 tiles, so the persistent grid has 6
    # workgroups, fewer than the 8 XCD groups; every tile must still be
    # computed (a partition over 8 groups would drop two panels)
    meshes, states, dts = _members(ccs, (24,), seed0=51)
    with ccs.Ensemble(meshes, dts, precision="fp32", tile_y=32, persistent=1) as ens:
        ens.set_state(0, states[0])
        ens.step(21)
        got = ens.get_state(0)
    monkeypatch.setenv("CCS_SPLIT_TILE_Y", "32")
    ref, _ = _single(ccs, meshes[0], states[0], dts[0], 21)
    assert _same(got, ref)


def test_autotune_choice_is_cached_per_member_sizes(ccs):
    sizes = (36, 28, 36)
    meshes, states, dts = _members(ccs, sizes, seed0=61)
    choices = []
    for _ in range(2):
        with ccs.Ensemble(meshes, dts, precision="fp32") as ens:
            for k, s in enumerate(states):
                ens.set_state(k, s)
            choices.append(ens.autotune())
    print(f"autotune: {choices}", flush=True)
    assert not choices[0]["from_cache"] and choices[1]["from_cache"]
    assert (choices[0]["tile_y"], choices[0]["persistent"]) == (choices[1]["tile_y"], choices[1]["persistent"])


def test_64_mixed_members_reach_a_common_simulated_time(ccs):
    # the benchmark ensemble: 64 members, N drawn from {48, 96, 144, 192},
    # each with its own CFL time step; run to one simulated hour, every member
    # at its own int(T / dt) steps, bit for bit its single run
    from conformal_cubed_sphere_amd.ensemble_run import build_members

    runs, ns = build_members(64)
    t_end = 3600.0
    with ccs.Ensemble([r[0] for r in runs], [r[2] for r in runs], precision="fp32") as ens:
        for k, r in enumerate(runs):
            ens.set_state(k, r[1])
        choice = ens.autotune()
        ens.run_until(t_end)
        ens.synchronize()
        want = [int(t_end / r[2]) for r in runs]
        print(f"64 members, autotune {choice}, steps {min(want)}..{max(want)}", flush=True)
        env = os.environ.get("CCS_SPLIT_TILE_Y")
        os.environ["CCS_SPLIT_TILE_Y"] = str(choice["tile_y"])
        try:
            for k, r in enumerate(runs):
                assert ens.steps_done(k) == want[k]
                ref, _ = _single(ccs, r[0], r[1], r[2], want[k])
                assert _same(ens.get_state(k), ref), k
        finally:
            if env is None:
                os.environ.pop("CCS_SPLIT_TILE_Y", None)
            else:
                os.environ["CCS_SPLIT_TILE_Y"] = env


@pytest.mark.parametrize("persistent", [0, 1])
def test_fp64_fast_members_match_the_oracle(ccs, orc, persistent):
    # the split kernel in FP64 evaluates the reference operator in a
    # non-dimensional form with shared smoothness indicators: equal to the
    # oracle (spatial_operators_mod.F90:13-181, time_scheme_mod.F90:46-67) to
    # FP64 round-off, not bit for bit
    sizes = (32, 48)
    meshes, states, dts = _members(ccs, sizes, seed0=71)
    nsteps = [int(6 * 3600.0 / dt) for dt in dts]
    with ccs.Ensemble(meshes, dts, precision="fp64_fast", tile_y=16, persistent=persistent) as ens:
        for k, s in enumerate(states):
            ens.set_state(k, s)
        ens.run_until(6 * 3600.0)
        for k in range(len(ens)):
            got = ens.get_state(k)
            o = orc.Oracle(meshes[k], states[k], dts[k])
            o.steps(nsteps[k])
            mesh = meshes[k]
            area = mesh.interior(mesh.areaCell)
            l2 = ccs.error_norms(mesh.interior(got.phi) / G, mesh.interior(o.a["phi"]) / G, area)[1]
            mask = ~mesh.corner_mask().ravel()
            worst = max(float(np.max(np.abs(getattr(got, f)[mask] - o.a[f][mask])) / np.max(np.abs(o.a[f][mask])))
                        for f in ("phiG", "u", "v"))
            print(f"member {k} (C{sizes[k]}, {nsteps[k]} steps): L2(h) {l2:.3e}, max rel diff {worst:.3e}", flush=True)
            assert l2 <= 1e-12
            assert worst <= 1e-11


def test_fp32_members_at_the_oracle_floor(ccs, orc):
    # FP32: the member's distance from the FP64 oracle stays within a small
    # factor of the distance between the oracle and the oracle started from
    # the state perturbed at FP32 rounding level (halo-consistent)
    sizes = (32, 48)
    meshes, states, dts = _members(ccs, sizes, seed0=81)
    t_end = 6 * 3600.0
    with ccs.Ensemble(meshes, dts, precision="fp32", tile_y=16, persistent=0) as ens:
        for k, s in enumerate(states):
            ens.set_state(k, s)
        ens.run_until(t_end)
        for k in range(len(ens)):
            mesh, nk = meshes[k], int(t_end / dts[k])
            got = ens.get_state(k)
            o = orc.Oracle(mesh, states[k], dts[k])
            o.steps(nk)
            op = orc.Oracle(mesh, ccs.perturb_state(mesh, states[k], 2.0 ** -24), dts[k])
            op.steps(nk)
            area = mesh.interior(mesh.areaCell)
            ref = mesh.interior(o.a["phi"]) / G
            e32 = ccs.error_norms(mesh.interior(got.phi) / G, ref, area)[1]
            efl = ccs.error_norms(mesh.interior(op.a["phi"]) / G, ref, area)[1]
            print(f"member {k} (C{sizes[k]}, {nk} steps): L2(h) fp32 {e32:.3e}, floor {efl:.3e}", flush=True)
            assert e32 <= 4.0 * efl + 1e-9
