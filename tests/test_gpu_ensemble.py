"""Batched ensemble (ccs_ensemble_*): mixed-resolution members advanced by one
launch per RK stage must evolve exactly as the same runs made one at a time."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = (16, 24, 32, 20)


def _members(ccs, sizes=SIZES, hours=6.0):
    meshes, states, dts = [], [], []
    for k, n in enumerate(sizes):
        m = ccs.Mesh(n, halo=3, radius=6.371e6)
        s = ccs.init_case(m, case_num=7, seed=11 + k, mountain=(k % 2 == 0))
        dt, _ = ccs.cfl_time_step(m, s, 0.5, hours * 3600.0)
        meshes.append(m)
        states.append(s)
        dts.append(dt)
    return meshes, states, dts


def _single(ccs, mesh, s0, dt, precision, scheme, nsteps):
    with ccs.DynamicalCore(mesh, dt, precision=precision, scheme=scheme) as core:
        core.set_state(s0)
        core.step(nsteps)
        return core.get_state(), core.diagnostics()


@pytest.mark.parametrize("precision,scheme,persistent", [
    ("fp32", "RK3_TVD", 0), ("fp32", "RK3_TVD", 1),
    ("fp32", "RK4", 1), ("fp32", "RK4", 0),
    ("fp64_fast", "RK3_TVD", 0), ("fp64_fast", "RK3_TVD", 1), ("fp64_fast", "RK4", 1)])
def test_members_equal_single_runs(ccs, precision, scheme, persistent):
    assert "CCS_SPLIT_TILE_Y" not in os.environ
    meshes, states, dts = _members(ccs)
    nsteps = 21  # one 16-step graph and five 1-step graphs
    print(f"{precision} {scheme} persistent={persistent}: members {SIZES}, {nsteps} steps", flush=True)
    with ccs.Ensemble(meshes, dts, precision=precision, scheme=scheme, tile_y=16, persistent=persistent) as ens:
        for k, s in enumerate(states):
            ens.set_state(k, s)
        ens.step(5)
        mid = [ens.diagnostics(k) for k in range(len(ens))]   # host work between graph replays
        ens.step(nsteps - 5)
        for k in range(len(ens)):
            got = ens.get_state(k)
            ref, (m1, e1) = _single(ccs, meshes[k], states[k], dts[k], precision, scheme, nsteps)
            for f in ("phiG", "u", "v"):
                assert np.array_equal(getattr(got, f), getattr(ref, f)), (k, f)
            assert ens.diagnostics(k) == (m1, e1)
            assert np.isfinite(mid[k][0])


def test_autotuned_members_equal_single_runs(ccs):
    # whatever tile height the autotuner picks, members equal the single runs
    # (which use 16-row tiles): results do not depend on the tile height
    assert "CCS_SPLIT_TILE_Y" not in os.environ
    meshes, states, dts = _members(ccs, sizes=(48, 24, 32))
    with ccs.Ensemble(meshes, dts, precision="fp32") as ens:
        for k, s in enumerate(states):
            ens.set_state(k, s)
        choice = ens.autotune()
        print(f"autotune: {choice}", flush=True)
        assert choice["tile_y"] in (8, 16, 32) and choice["persistent"] in (0, 1) and choice["stage_ms"] > 0
        ens.step(17)
        got = [ens.get_state(k) for k in range(len(ens))]
    for k in range(len(meshes)):
        ref, _ = _single(ccs, meshes[k], states[k], dts[k], "fp32", "RK3_TVD", 17)
        for f in ("phiG", "u", "v"):
            assert np.array_equal(getattr(got[k], f), getattr(ref, f)), (k, f)


@pytest.mark.parametrize("tile_y,persistent", [(8, 0), (32, 0), (8, 1), (32, 1)])
def test_members_at_every_tile_height_equal_single_runs(ccs, tile_y, persistent):
    assert "CCS_SPLIT_TILE_Y" not in os.environ
    meshes, states, dts = _members(ccs, sizes=(40, 24, 100))
    print(f"fp32 ensemble, {tile_y}-row tiles, persistent={persistent}: 18 steps", flush=True)
    with ccs.Ensemble(meshes, dts, precision="fp32", tile_y=tile_y, persistent=persistent) as ens:
        for k, s in enumerate(states):
            ens.set_state(k, s)
        ens.step(18)
        got = [ens.get_state(k) for k in range(len(ens))]
    for k in range(len(meshes)):
        ref, _ = _single(ccs, meshes[k], states[k], dts[k], "fp32", "RK3_TVD", 18)
        for f in ("phiG", "u", "v"):
            assert np.array_equal(getattr(got[k], f), getattr(ref, f)), (k, f)


@pytest.mark.parametrize("persistent", [0, 1])
def test_members_at_the_benchmark_sizes_equal_single_runs(ccs, persistent):
    # one member of each size class of the 64-member benchmark ensemble
    # (N in {48, 96, 144, 192}), 21 steps: one 16-step graph, five single steps
    sizes = (48, 96, 144, 192)
    meshes, states, dts = _members(ccs, sizes=sizes)
    print(f"fp32 RK3_TVD persistent={persistent}: members {sizes}, 21 steps", flush=True)
    with ccs.Ensemble(meshes, dts, precision="fp32", tile_y=16, persistent=persistent) as ens:
        for k, s in enumerate(states):
            ens.set_state(k, s)
        ens.step(21)
        for k in range(len(ens)):
            got = ens.get_state(k)
            ref, (m1, e1) = _single(ccs, meshes[k], states[k], dts[k], "fp32", "RK3_TVD", 21)
            for f in ("phiG", "u", "v"):
                assert np.array_equal(getattr(got, f), getattr(ref, f)), (k, f)
            assert ens.diagnostics(k) == (m1, e1)


def test_member_mass_is_conserved(ccs):
    meshes, states, dts = _members(ccs, sizes=(24, 32), hours=24.0)
    with ccs.Ensemble(meshes, dts, precision="fp32", tile_y=16, persistent=1) as ens:
        for k, s in enumerate(states):
            ens.set_state(k, s)
        m0 = [ens.diagnostics(k)[0] for k in range(2)]
        ens.step(100)
        for k in range(2):
            m1 = ens.diagnostics(k)[0]
            print(f"member {k}: mass drift {(m1 - m0[k]) / m0[k]:.3e}", flush=True)
            assert abs(m1 - m0[k]) <= 1e-9 * m0[k]


def test_step_before_state_is_an_order_error(ccs):
    meshes, states, dts = _members(ccs, sizes=(16, 16))
    with ccs.Ensemble(meshes, dts) as ens:
        ens.set_state(0, states[0])
        with pytest.raises(ccs.CcsError) as ei:
            ens.step(1)
        assert ei.value.code == 3
