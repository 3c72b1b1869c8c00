"""Accuracy of the split-precision path against the FP64 GPU path.

``python -m conformal_cubed_sphere_amd.accuracy_run --n 720 --days 5``

Runs one initial state (SSP-RK3, CFL 0.5, a = 6.371e6 m) several times:
- FP64 (reference arithmetic: the parity kernel, or the split kernel in FP64);
- FP32 split precision;
- FP64 started from the state perturbed by perturb_state at relative sizes
  2^-k for each k of --perturb (default 24: one FP32 rounding).  The
  perturbation is halo-consistent: every halo cell is perturbed exactly as
  the interior cell it copies.  The distance of such a run from the FP64 run
  is the method's own sensitivity to perturbations of that size.

The initial state is the seeded balanced jet with 8 bumps and a Gaussian
mountain (--case 7, the benchmark recipe) or Williamson case 2, the steady
zonal flow of test_case_mod.F90:61-95 (--case 2), whose exact solution does
not change in time, so rounding errors are not amplified by developing
structure.

Every run records, at --snapshots evenly spaced times (or at the simulated
--hours given), the normalised L2 of h
against the FP64 run at the same time; at the end the normalised L1/L2/Linf of
h, zonal u and meridional v (area-weighted, as doc/orderCheck.m:55-74).  One
JSON line is printed at the end; progress lines go to stderr.  `--only fp32`
runs the FP32 path alone and reports its drifts, throughput and the range of h
(for runs too long to repeat in FP64).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .core import GRAVITY, CcsError, DynamicalCore, Mesh, cfl_time_step, error_norms, init_case, perturb_state


def _run(mesh, s0, dt, nsteps, precision, with_companions, chunk, label, marks=(), ref_h=None):
    """Run nsteps; at each step count in marks record interior h (ref_h None)
    or its normalised L2 against ref_h[mark]."""
    t0 = time.perf_counter()
    area = mesh.interior(mesh.areaCell)
    series = {}
    with DynamicalCore(mesh, dt, precision=precision) as core:
        core.set_state(s0, with_companions=with_companions)
        m0, e0 = core.diagnostics()
        done = 0
        stops = sorted(set(list(marks) + [nsteps]))
        for stop in stops:
            while done < stop:
                k = min(chunk, stop - done)
                core.step(k)
                core.synchronize()
                done += k
                print(f"{label}: {done}/{nsteps} steps, {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                      flush=True)
            if stop in marks:
                h = mesh.interior(core.get_state().phi) / GRAVITY
                if ref_h is None:
                    series[stop] = h.copy()
                else:
                    series[stop] = error_norms(h, ref_h[stop], area)[1]
        try:
            m1, e1 = core.diagnostics()
        except CcsError:  # the state went non-finite: the drift and the norms come out NaN
            m1 = e1 = float("nan")
        st = core.get_state()
    wall = time.perf_counter() - t0
    return st, series, {"mass_drift": (m1 - m0) / m0, "energy_drift": (e1 - e0) / e0,
                        "cell_updates_per_s": 6 * mesh.n * mesh.n * nsteps / wall, "wall_s": wall}


def _drift_series(mesh, s0, dt, precision, marks):
    """Relative mass and energy drift after each step count in marks."""
    out = {"mass": [], "energy": []}
    with DynamicalCore(mesh, dt, precision=precision) as core:
        core.set_state(s0)
        m0, e0 = core.diagnostics()
        done = 0
        for m in marks:
            core.step(m - done)
            done = m
            try:
                m1, e1 = core.diagnostics()
            except CcsError:  # non-finite state
                m1 = e1 = float("nan")
            out["mass"].append((m1 - m0) / m0)
            out["energy"].append((e1 - e0) / e0)
    return out


def _survival(mesh, s0, dt, nsteps, precision, chunk, step0=0, record=False):
    """Step from s0 in chunks, checking that h and the zonal wind stay finite;
    the largest |u| and smallest h seen, and when the state went non-finite."""
    out = {"blew_up_after_step": None, "blew_up_by_hours": None, "umax": 0.0, "hmin": float("inf")}
    rows = []
    with DynamicalCore(mesh, dt, precision=precision) as core:
        core.set_state(s0)
        done = 0
        while done < nsteps:
            k = min(chunk, nsteps - done)
            core.step(k)
            done += k
            st = core.get_state()
            h = mesh.interior(st.phi) / GRAVITY
            u = np.abs(mesh.interior(st.zonal))
            finite = bool(np.all(np.isfinite(h)) and np.all(np.isfinite(u)))
            if record:
                rows.append({"step": step0 + done, "finite": finite, "umax": float(np.nanmax(u)),
                             "hmin": float(np.nanmin(h))})
            if not finite:
                out["blew_up_after_step"] = step0 + done
                out["blew_up_by_hours"] = (step0 + done) * dt / 3600.0
                break
            out["umax"] = max(out["umax"], float(np.max(u)))
            out["hmin"] = min(out["hmin"], float(np.min(h)))
    if record:
        out["rows"] = rows
    return out


def _norms(mesh, st, ref):
    area = mesh.interior(mesh.areaCell)
    out = {}
    for name, f in (("h", lambda s: s.phi / GRAVITY), ("u", lambda s: s.zonal), ("v", lambda s: s.merid)):
        l1, l2, linf = error_norms(mesh.interior(f(st)), mesh.interior(f(ref)), area)
        out[name] = {"L1": l1, "L2": l2, "Linf": linf}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=720)
    ap.add_argument("--days", type=float, default=5.0)
    ap.add_argument("--case", type=int, default=7, choices=[2, 7],
                    help="7: seeded jet + bumps + mountain (benchmark recipe); 2: Williamson steady zonal flow")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--reference", default="fp64", choices=["fp64", "fp64_fast"],
                    help="FP64 reference: the parity kernel or the split kernel instantiated in FP64")
    ap.add_argument("--perturb", default="24",
                    help="comma-separated k: FP64 runs from the state perturbed by 2^-k (empty: none)")
    ap.add_argument("--snapshots", type=int, default=0, help="L2(h) against FP64 at this many evenly spaced times")
    ap.add_argument("--hours", default="", help="comma-separated simulated hours for the L2(h) series instead")
    ap.add_argument("--chunk", type=int, default=4000)
    ap.add_argument("--only", choices=["fp32"], default=None,
                    help="run only the FP32 path (drift, stability and throughput; no error norms)")
    ap.add_argument("--drift-steps", default="",
                    help="comma-separated step counts: relative mass and energy drift of FP32, FP64 parity, "
                         "FP64 from the 2^-24-perturbed state and the split kernel in FP64 at each (no error norms)")
    ap.add_argument("--cfl", type=float, default=0.5, help="Courant number of the time step (initial state)")
    ap.add_argument("--survival", default="",
                    help="comma-separated k (0: unperturbed): does the run from the state perturbed by 2^-k stay "
                         "finite for --days (checked every --chunk steps)?  Precision from --precision")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64_fast", "fp64"],
                    help="precision of the --survival runs")
    ap.add_argument("--branch-at", type=int, default=0,
                    help="run FP32 this many steps, then continue that state in FP32, FP64-fast and FP64 parity "
                         "for --branch-steps steps, checking every 50 steps")
    ap.add_argument("--branch-steps", type=int, default=400)
    args = ap.parse_args()

    mesh = Mesh(args.n, halo=3, radius=6.371e6, omega=7.292e-5)
    if args.case == 7:
        s0 = init_case(mesh, case_num=7, seed=args.seed, n_bumps=8, mountain=True, gravity=GRAVITY)
    else:
        s0 = init_case(mesh, case_num=2, gravity=GRAVITY)
    dt, nsteps = cfl_time_step(mesh, s0, args.cfl, args.days * 86400.0)
    if args.survival:
        out = {"n": args.n, "case": args.case, "days": args.days, "cfl": args.cfl, "dt_s": dt, "steps": nsteps,
               "precision": args.precision, "runs": []}
        for k in [int(x) for x in args.survival.split(",")]:
            st = s0 if k == 0 else perturb_state(mesh, s0, 2.0 ** -k)
            out["runs"].append({"perturbation": f"2^-{k}" if k else "none",
                                **_survival(mesh, st, dt, nsteps, args.precision, args.chunk)})
            print(json.dumps(out["runs"][-1]), file=sys.stderr, flush=True)
        print(json.dumps(out), flush=True)
        return
    if args.branch_at:
        out = {"n": args.n, "case": args.case, "cfl": args.cfl, "dt_s": dt, "branch_at": args.branch_at,
               "continued": {}}
        with DynamicalCore(mesh, dt, precision="fp32") as core:
            core.set_state(s0)
            core.step(args.branch_at)
            sb = core.get_state()
        for prec in ("fp32", "fp64_fast", "fp64"):
            out["continued"][prec] = _survival(mesh, sb, dt, args.branch_steps, prec, 50, args.branch_at,
                                               record=True)
        print(json.dumps(out), flush=True)
        return
    if args.drift_steps:
        marks = sorted({int(x) for x in args.drift_steps.split(",")})
        out = {"n": args.n, "case": args.case, "dt_s": dt, "steps": marks, "drift": {}}
        for name, prec, st in (("fp32", "fp32", s0), ("fp64", "fp64", s0),
                               ("fp64_perturbed_2^-24", "fp64", perturb_state(mesh, s0, 2.0 ** -24)),
                               ("fp64_fast", "fp64_fast", s0)):
            out["drift"][name] = _drift_series(mesh, st, dt, prec, marks)
        print(json.dumps(out), flush=True)
        return
    if args.only == "fp32":
        f32, _, f32_info = _run(mesh, s0, dt, nsteps, "fp32", True, args.chunk, "fp32")
        h = mesh.interior(f32.phi) / GRAVITY
        out = {"n": args.n, "case": args.case, "days": args.days, "steps": nsteps, "dt_s": dt, "fp32": f32_info,
               "finite": bool(np.all(np.isfinite(h))), "h_min_m": float(np.min(h)), "h_max_m": float(np.max(h))}
        print(json.dumps(out), flush=True)
        return
    marks = [round(nsteps * (k + 1) / args.snapshots) for k in range(args.snapshots)] if args.snapshots else []
    if args.hours:
        marks = sorted({min(nsteps, max(1, round(float(x) * 3600.0 / dt))) for x in args.hours.split(",")})
    ref, ref_h, ref_info = _run(mesh, s0, dt, nsteps, args.reference, True, args.chunk, args.reference, marks)
    f32, f32_l2, f32_info = _run(mesh, s0, dt, nsteps, "fp32", True, args.chunk, "fp32", marks, ref_h)
    out = {"n": args.n, "case": args.case, "days": args.days, "steps": nsteps, "dt_s": dt,
           "reference": args.reference, "snapshot_hours": [m * dt / 3600.0 for m in marks],
           "fp32_vs_fp64": _norms(mesh, f32, ref), "fp32_l2_h_series": [f32_l2[m] for m in marks],
           "fp32": f32_info, args.reference: ref_info, "perturbed": {},
           "fp32_finite": bool(np.all(np.isfinite(mesh.interior(f32.phi))))}
    for k in [int(x) for x in args.perturb.split(",") if x.strip()]:
        sp = perturb_state(mesh, s0, 2.0 ** -k)
        pert, pert_l2, pert_info = _run(mesh, sp, dt, nsteps, args.reference, True, args.chunk,
                                        f"perturbed 2^-{k}", marks, ref_h)
        out["perturbed"][f"2^-{k}"] = {"vs_fp64": _norms(mesh, pert, ref), "l2_h_series": [pert_l2[m] for m in marks],
                                       **pert_info}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
