#!/usr/bin/env python3
"""Throughput of the dynamic-core time step on one MI355X.

Workload (the configuration the headline metric is quoted on): C720
(6 x 720^2 cells, halo 3), seeded balanced jet + 8 Gaussian bumps + Gaussian
mountain, SSP-RK3 at CFL 0.5 (dt from the initial state over a 5-day run),
a = 6.371e6 m, g = 9.80616, Omega = 7.292e-5.  A step is one SSP-RK3 step of
all six panels (3 fused stage kernels).  Inputs are resident in HBM before the
timed region.  The split-precision path (FP32 state) is the headline `value`;
the FP64 parity path is timed beside it at the same configuration.

Before the W warmup steps every timed path runs untimed steps for 0.3 s
(`settle_steps` in the line): the GPU's clocks ramp up over the first ~0.1 s
of sustained work, and a 20-step timed region would otherwise be measured on
a GPU that is still ramping.

Prints one JSON line.  `python bench.py --gpus N --steps K --warmup W`.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s at C720, FP32 vs FP64 GPU at matched L2(h) error; HBM GB/s"
HBM_PEAK_GBPS = 8000.0
GRAVITY = 9.80616
WORD = {"fp32": 4, "fp64_fast": 8, "fp64": 8}
# Algorithmic bytes of one cell-update (one interior cell, one SSP-RK3 step),
# the unit of the roofline figure: 24 state words (3 stages x (3 read + 3
# written) + 2 stages x 3 base-state reads) and 9 static words (3 stages x
# {sqrtG, f, phis}; G^-1 = I / sqrtG on the conformal grid): 33 words, 132 B
# in FP32, 264 B in FP64.  Per stage launch: a third of that.
UPDATE_WORDS = 33
# What each stage kernel reads and writes besides (mean words per cell per
# launch).  Split path: stage input (dP, U, V) + statics {Pref, sqrtg, phis,
# a sqrtg f} + output (3), + base state in stages 2 and 3: (10 + 13 + 13) / 3
# = 12 (the reference field Pref of the FP32 deviation is the word beyond the
# 11 of the cell-update count).  FP64 parity path: input (3) + statics
# {sqrtG, G^-1 (4), phis, f, A G^-1 (4), G A^-1 (4)} + output (3), + base (3)
# in stages 2 and 3: (21 + 24 + 24) / 3 = 23.
KERNEL_WORDS_PER_LAUNCH = {"split": 12.0, "parity": 23.0}


def update_bytes_per_launch(precision):
    """Algorithmic bytes per interior cell per stage launch (cell-update bytes / 3)."""
    return WORD[precision] * UPDATE_WORDS / 3.0


def kernel_bytes_per_launch(precision):
    """Bytes per interior cell per stage launch that the kernel as built moves."""
    return WORD[precision] * KERNEL_WORDS_PER_LAUNCH["parity" if precision == "fp64" else "split"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=720)
    ap.add_argument("--precision", default="auto", choices=["auto", "fp32", "fp64"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ens-members", type=int, default=64)
    ap.add_argument("--ens-steps", type=int, default=64)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--spatial-op", default="weno",
                    choices=["weno", "centred2", "centred4", "centred6", "weno_ext"],
                    help="spatial_operators_mod variant (weno_ext runs on the FP64 parity path only)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device copies of MIB MiB (--reps times), the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE to HBM bytes, and exit")
    ap.add_argument("--reps", type=int, default=5, help="copies of --copy-calibration")
    return ap.parse_args()


def copy_calibration(mib, reps):
    """A device-to-device copy of exactly mib MiB read and mib MiB written per
    dispatch (torch as device-memory plumbing); under rocprofv3 --pmc the
    counters it reads give the scale from counter units to HBM bytes."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x {reps}: read {mib} MiB and wrote {mib} MiB per dispatch", flush=True)


def time_path(ccs, mesh, s0, dt, precision, steps, warmup, spatial_op="weno", settle_s=0.3):
    core = ccs.DynamicalCore(mesh, dt, precision=precision, use_graph=True, spatial_op=spatial_op)
    core.set_state(s0)
    core.step(1)          # first step reads the uploaded companions
    # untimed steps until the GPU has run for settle_s (its clocks ramp up
    # over the first ~0.1 s of sustained work), then the warmup steps
    settle = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < settle_s:
        core.step(16)
        core.synchronize()
        settle += 16
    core.step(warmup)
    core.synchronize()
    t0 = time.perf_counter()
    core.step(steps)
    core.synchronize()
    wall = time.perf_counter() - t0
    ev_total, _ = core.time_steps(steps)            # HIP events around the graph replays
    nk = min(steps, 60)
    _, kern_ms = core.time_steps(nk, kernel_timing=True)  # events attached to every stage dispatch
    core.advanced = 1 + settle + warmup + 2 * steps + 2 * nk  # time_steps(kernel_timing) runs two passes
    core.settle_steps = settle
    return core, wall, ev_total, kern_ms


def field_norms(ccs, mesh, st, ref):
    """Normalised, area-weighted L1/L2/Linf of h, zonal and meridional wind
    of state st against ref (interior cells)."""
    area = mesh.interior(mesh.areaCell)
    out = {}
    for name, f in (("h", lambda s: s.phi / GRAVITY), ("u", lambda s: s.zonal), ("v", lambda s: s.merid)):
        l1, l2, linf = ccs.error_norms(mesh.interior(f(st)), mesh.interior(f(ref)), area)
        out[name] = {"L1": l1, "L2": l2, "Linf": linf}
    return out


def drift(d0, d1):
    """Relative mass and energy drift from diagnostics (mass, energy) d0 to d1."""
    return {"mass": (d1[0] - d0[0]) / d0[0], "energy": (d1[1] - d0[1]) / d0[1]}


def host_cores():
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return int(os.environ.get("OMP_NUM_THREADS", cores))


def cpu_baseline_reference(n, budget_s):
    """The reference's own step (oracle/_ref/libccs_ref_omp.so: its modules
    built from source with its GNU flags, OpenMP on), timed in a child
    process by oracle/ref_timing.py; None when that library is not built."""
    import subprocess

    if not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libccs_ref_omp.so")):
        return None
    cores = host_cores()
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    try:
        r = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "ref_timing.py"), str(n), str(budget_s)],
                           capture_output=True, text=True, timeout=30 * budget_s + 120, env=env)
        line = [x for x in r.stdout.splitlines() if x.startswith("{")]
        res = json.loads(line[-1]) if line else {"error": r.stderr[-300:]}
    except (subprocess.TimeoutExpired, ValueError) as exc:
        res = {"error": repr(exc)}
    if "error" in res:
        return None
    return {"value": res["value"], "unit": "cell-updates/s", "cores": cores, "kind": "reference",
            "sample": f"the reference's dynamic-core modules (spatial_operators_mod, ghost_mod, projection_mod, "
                      f"... compiled from source with -O3 -fdefault-real-8 -fopenmp) in an RK3_TVD loop "
                      f"(oracle/ref_harness.F90, oracle/ref_timing.py) on the same C{n} workload: "
                      f"{res['steps']} SSP-RK3 steps after 1 untimed step, {res['seconds']:.1f} s on {cores} "
                      f"OpenMP threads"}


def cpu_baseline(ccs, mesh, s0, dt, budget_s, spatial_op="weno"):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib  # the checker, timed here as the CPU baseline

    cores = host_cores()
    orc = oracle_lib.Oracle(mesh, s0, dt, threads=cores, op=spatial_op)
    t0 = time.perf_counter()
    orc.steps(1)
    one = time.perf_counter() - t0
    k = max(1, min(50, int(budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    orc.steps(k)
    el = time.perf_counter() - t0
    cells = 6 * mesh.n * mesh.n
    return {"value": cells * k / el, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": f"oracle/ccs_oracle.c (FP64 restatement of the reference step, OpenMP over rows) "
                      f"on the same C{mesh.n} workload: {k} SSP-RK3 steps after 1 untimed step, "
                      f"{el:.1f} s on {cores} host threads"}


def ensemble_bench(ccs, members, steps, warmup, settle_s=0.3):
    """64 independent runs, N drawn from {48, 96, 144, 192}, per-member seeds
    for bumps and mountain, CFL 0.5 steps of their own: batched (one launch
    per RK stage for all members, autotuned) against the same runs made one
    after another; per-member error against the FP64 parity path."""
    import numpy as np

    from conformal_cubed_sphere_amd.ensemble_run import build_members

    runs, sizes = build_members(members)
    cells = sum(6 * n * n for n in sizes)

    ens = ccs.Ensemble([r[0] for r in runs], [r[2] for r in runs], precision="fp32")
    for k, r in enumerate(runs):
        ens.set_state(k, r[1])
    choice = ens.autotune()
    ens.step(warmup)  # captures the graphs the timed steps replay
    ens.synchronize()
    # the GPU's clocks ramp up over the first tens of milliseconds of
    # sustained work and drop while it idles (as during uploads and graph
    # captures): untimed batched steps of a second copy of the ensemble for
    # settle_s, then the timed steps at once.  The second copy runs the other
    # launch grid at the chosen tile height (the persistent grid, whose
    # workgroups claim tiles from per-XCD counters, or one workgroup per
    # tile), and is timed after the batched steps on the same warm clocks.
    other = {"tile_y": choice["tile_y"], "persistent": 1 - choice["persistent"]}
    warm = ccs.Ensemble([r[0] for r in runs], [r[2] for r in runs], precision="fp32", **other)
    for k, r in enumerate(runs):
        warm.set_state(k, r[1])
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < settle_s:
        warm.step(16)
        warm.synchronize()
    t0 = time.perf_counter()
    ens.step(steps)
    ens.synchronize()
    wall_b = time.perf_counter() - t0
    t0 = time.perf_counter()
    warm.step(steps)
    warm.synchronize()
    other["value"] = cells * steps / (time.perf_counter() - t0)
    warm.close()

    cores = []
    for r in runs:
        c = ccs.DynamicalCore(r[0], r[2], precision="fp32")
        c.set_state(r[1])
        c.step(warmup)
        c.synchronize()
        cores.append(c)
    for c in cores:  # an untimed pass of the same launches settles the clocks
        c.step(steps)
        c.synchronize()
    t0 = time.perf_counter()
    for c in cores:  # one run after another
        c.step(steps)
        c.synchronize()
    wall_s = time.perf_counter() - t0
    for c in cores:
        c.close()

    # error against FP64, beside the method's own floor: an FP64 run from the
    # initial state perturbed at FP32 rounding level (2^-24 relative)
    l2, floor, drift = [], [], []
    for k, r in enumerate(runs):
        got = ens.get_state(k)
        sp = ccs.perturb_state(r[0], r[1], 2.0 ** -24)
        fin = []
        for st0, first in ((r[1], True), (sp, False)):
            with ccs.DynamicalCore(r[0], r[2], precision="fp64") as c:
                c.set_state(st0)
                if first:
                    m0, _ = c.diagnostics()
                c.step(warmup + steps)
                fin.append(c.get_state())
        mesh = r[0]
        area = mesh.interior(mesh.areaCell)
        href = mesh.interior(fin[0].phi) / GRAVITY
        l2.append(ccs.error_norms(mesh.interior(got.phi) / GRAVITY, href, area)[1])
        floor.append(ccs.error_norms(mesh.interior(fin[1].phi) / GRAVITY, href, area)[1])
        drift.append(abs(ens.diagnostics(k)[0] - m0) / m0)
    ens.close()
    counts = {str(n): sizes.count(n) for n in sorted(set(sizes))}
    return {"members": members, "members_per_n": counts, "interior_cells": cells, "steps": steps,
            "value_batched": cells * steps / wall_b, "value_serial": cells * steps / wall_s,
            "unit": "member cell-updates/s", "speedup_batched_over_serial": wall_s / wall_b,
            "autotune": choice, "other_launch_grid": other, "precision": "fp32",
            "l2_h_vs_fp64_median": float(np.median(l2)), "l2_h_vs_fp64_max": float(np.max(l2)),
            "l2_h_fp64_perturbed_median": float(np.median(floor)), "l2_h_fp64_perturbed_max": float(np.max(floor)),
            "mass_drift_max": float(np.max(drift)),
            "note": f"error after {warmup + steps} steps of each member's own dt"}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration, args.reps)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); "
              "the multi-GPU panel sharding is a design only, not a runnable path", file=sys.stderr)
        sys.exit(2)
    import conformal_cubed_sphere_amd as ccs

    mesh = ccs.Mesh(args.n, halo=3, radius=6.371e6, omega=7.292e-5)
    s0 = ccs.init_case(mesh, case_num=7, seed=1, n_bumps=8, mountain=True, gravity=GRAVITY)
    dt, nsteps_5d = ccs.cfl_time_step(mesh, s0, 0.5, 5 * 86400.0)
    cells = 6 * args.n * args.n

    headline = args.precision
    if args.spatial_op == "weno_ext":
        headline = "fp64"  # the extended-lattice operator runs on the FP64 parity path only
    if headline == "auto":
        headline = "fp32"
        try:
            ccs.DynamicalCore(mesh, dt, precision="fp32").close()
        except ccs.CcsError:
            headline = "fp64"
    dtype = "f32" if headline == "fp32" else "f64"
    core, wall, ev_total, kern_ms = time_path(ccs, mesh, s0, dt, headline, args.steps, args.warmup,
                                              args.spatial_op)
    value = cells * args.steps / wall
    bpl = update_bytes_per_launch(headline)
    kbpl = kernel_bytes_per_launch(headline)
    achieved = bpl * cells / (kern_ms * 1e-3) / 1e9  # GB/s of the stage kernel
    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
        "config": {"workload": f"C{args.n} seeded balanced jet + 8 bumps + Gaussian mountain, SSP-RK3, "
                               f"CFL 0.5, halo 3, precision {headline}"
                               + ("" if args.spatial_op == "weno" else f", spatial operator {args.spatial_op}"),
                   "n": args.n, "halo": 3, "dt_s": dt, "steps_per_5_days": nsteps_5d,
                   "interior_cells": cells, "precision": headline, "spatial_op": args.spatial_op},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "kernel": "stage kernel (fused tendency + RK stage + halo exchange)",
                     "kernel_ms_avg": kern_ms, "algorithmic_bytes_per_launch": bpl * cells,
                     "algorithmic_bytes_per_cell_per_launch": bpl,
                     "algorithmic_bytes_per_cell_update": 3 * bpl,
                     "kernel_bytes_per_cell_per_launch": kbpl,
                     "achieved_at_kernel_bytes": kbpl * cells / (kern_ms * 1e-3) / 1e9,
                     "frac_at_kernel_bytes": kbpl * cells / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS,
                     "limiter": "hbm" if args.spatial_op.startswith("centred") else "valu",
                     "steps_per_s_over_stage_kernels": (3 * kern_ms) / (wall * 1e3 / args.steps)},
        "event_ms_per_step": ev_total / args.steps,
        "settle_steps": core.settle_steps,
    }
    prof = os.path.join(ROOT, "profiles", f"traffic_c{args.n}_{headline}.json")
    if os.path.exists(prof) and args.spatial_op == "weno":
        with open(prof) as f:
            t = json.load(f)
        out["roofline"]["traffic"] = t.get("hbm_bytes_per_launch")
        out["roofline"]["traffic_source"] = os.path.relpath(prof, ROOT)
    # the stage kernel is limited by VALU issue and latency, not by HBM: the
    # committed SQ counters of the same kernel say how much
    valu = os.path.join(ROOT, "profiles", f"valu_c{args.n}_{headline}.json")
    if os.path.exists(valu) and args.spatial_op == "weno":
        with open(valu) as f:
            v = json.load(f)
        out["roofline"]["valu"] = {"insts_per_wave": v.get("valu_insts_per_wave"),
                                   "wait_inst_any_frac": v.get("wait_inst_any_frac_of_wave_cycles"),
                                   "source": os.path.relpath(valu, ROOT)}

    if headline == "fp32" and not args.no_fp64:
        steps64 = max(20, args.steps // 3)
        c64, wall64, _, kern64 = time_path(ccs, mesh, s0, dt, "fp64", steps64, 5, args.spatial_op)
        v64 = cells * steps64 / wall64
        out["fp64_parity_path"] = {"value": v64, "unit": "cell-updates/s", "kernel_ms_avg": kern64,
                                   "hbm_gbps_algorithmic": update_bytes_per_launch("fp64") * cells / (kern64 * 1e-3) / 1e9,
                                   "hbm_gbps_kernel_bytes": kernel_bytes_per_launch("fp64") * cells / (kern64 * 1e-3) / 1e9}
        c64.close()
        cf, wallf, _, kernf = time_path(ccs, mesh, s0, dt, "fp64_fast", steps64, 5, args.spatial_op)
        out["fp64_split_kernel"] = {"value": cells * steps64 / wallf, "unit": "cell-updates/s", "kernel_ms_avg": kernf}
        cf.close()
        out["fp32_over_fp64"] = value / v64
        # Matched error: FP32 vs FP64 after the same steps from the same start,
        # beside the method's own floor (an FP64 run whose initial state is
        # perturbed at FP32 rounding level, 2^-24 relative, halo-consistent).
        s32 = core.get_state()
        d32 = core.diagnostics()
        core.close()
        nadv = core.advanced
        # the FP32 handle's own diagnostics of the uploaded state (its winds
        # are the FP32 casts) are the start of its drift
        with ccs.DynamicalCore(mesh, dt, precision="fp32", spatial_op=args.spatial_op) as c:
            c.set_state(s0)
            d32_0 = c.diagnostics()
        sp = ccs.perturb_state(mesh, s0, 2.0 ** -24)
        states, drift64 = {}, None
        for name, st0 in (("ref", s0), ("pert", sp)):
            c = ccs.DynamicalCore(mesh, dt, precision="fp64", spatial_op=args.spatial_op)
            c.set_state(st0)
            d0 = c.diagnostics()
            c.step(nadv)
            states[name] = c.get_state()
            if name == "ref":
                drift64 = drift(d0, c.diagnostics())
            c.close()
        area = mesh.interior(mesh.areaCell)
        href = mesh.interior(states["ref"].phi) / GRAVITY
        out["l2_h_fp32_vs_fp64"] = ccs.error_norms(mesh.interior(s32.phi) / GRAVITY, href, area)[1]
        out["l2_h_fp64_perturbed_vs_fp64"] = ccs.error_norms(mesh.interior(states["pert"].phi) / GRAVITY, href,
                                                             area)[1]
        out["l2_h_compared_after_steps"] = nadv
        out["l2_h_compared_after_hours"] = nadv * dt / 3600.0
        # normalised, area-weighted L1/L2/Linf of h, zonal and meridional
        # wind against the FP64 parity run, and relative mass/energy drift
        out["errors_vs_fp64"] = {"fp32": field_norms(ccs, mesh, s32, states["ref"]),
                                 "fp64_perturbed_2^-24": field_norms(ccs, mesh, states["pert"], states["ref"])}
        out["drift"] = {"fp32": drift(d32_0, d32), "fp64": drift64}
    else:
        core.close()
    if not args.no_ensemble:
        out["ensemble"] = ensemble_bench(ccs, args.ens_members, args.ens_steps, 16)
    if not args.no_cpu:
        port = cpu_baseline(ccs, mesh, s0, dt, args.cpu_seconds, args.spatial_op)
        # the timed reference build holds the default operator module
        ref = cpu_baseline_reference(args.n, args.cpu_seconds) if args.spatial_op == "weno" else None
        out["cpu_baseline"] = ref if ref is not None else port
        if ref is not None:
            out["cpu_oracle_port"] = port  # the C restatement, for comparison
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
