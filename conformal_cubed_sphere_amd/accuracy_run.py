"""Accuracy of the split-precision path against the FP64 GPU path.

``python -m conformal_cubed_sphere_amd.accuracy_run --n 720 --days 5``

Runs the seeded balanced jet with 8 bumps and a Gaussian mountain (SSP-RK3,
CFL 0.5, a = 6.371e6 m) three times from the same start:
- FP64 (reference arithmetic);
- FP32 split precision;
- FP64 started from a state perturbed at FP32 rounding level, 2^-24 relative.
  Its distance from the FP64 run is the method's own sensitivity floor.

Afterwards it prints one JSON line with:
- the normalised L1/L2/Linf of h, zonal u and meridional v (area-weighted, as
  doc/orderCheck.m:55-74) for FP32 and for the perturbed run, both against FP64;
- the relative mass and energy drift of each run;
- the throughput of each run.

Progress lines go to stderr.  `--only fp32` runs the FP32 path alone and
reports its drifts, throughput and the range of h (for runs too long to
repeat in FP64).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .core import GRAVITY, DynamicalCore, Mesh, cfl_time_step, error_norms, init_case


def _run(mesh, s0, dt, nsteps, precision, with_companions, chunk, label, fp64_ref=None):
    t0 = time.perf_counter()
    with DynamicalCore(mesh, dt, precision=precision) as core:
        core.set_state(s0, with_companions=with_companions)
        m0, e0 = core.diagnostics()
        done = 0
        while done < nsteps:
            k = min(chunk, nsteps - done)
            core.step(k)
            core.synchronize()
            done += k
            print(f"{label}: {done}/{nsteps} steps, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        m1, e1 = core.diagnostics()
        st = core.get_state()
    wall = time.perf_counter() - t0
    return st, {"mass_drift": (m1 - m0) / m0, "energy_drift": (e1 - e0) / e0,
                "cell_updates_per_s": 6 * mesh.n * mesh.n * nsteps / wall, "wall_s": wall}


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
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--reference", default="fp64", choices=["fp64", "fp64_fast"],
                    help="FP64 reference: the parity kernel or the split kernel instantiated in FP64")
    ap.add_argument("--no-floor", action="store_true", help="skip the perturbed FP64 run")
    ap.add_argument("--chunk", type=int, default=4000)
    ap.add_argument("--only", choices=["fp32"], default=None,
                    help="run only the FP32 path (drift, stability and throughput; no error norms)")
    args = ap.parse_args()

    mesh = Mesh(args.n, halo=3, radius=6.371e6, omega=7.292e-5)
    s0 = init_case(mesh, case_num=7, seed=args.seed, n_bumps=8, mountain=True, gravity=GRAVITY)
    dt, nsteps = cfl_time_step(mesh, s0, 0.5, args.days * 86400.0)
    if args.only == "fp32":
        f32, f32_info = _run(mesh, s0, dt, nsteps, "fp32", True, args.chunk, "fp32")
        h = mesh.interior(f32.phi) / GRAVITY
        out = {"n": args.n, "days": args.days, "steps": nsteps, "dt_s": dt, "fp32": f32_info,
               "finite": bool(np.all(np.isfinite(h))), "h_min_m": float(np.min(h)), "h_max_m": float(np.max(h))}
        print(json.dumps(out), flush=True)
        return
    ref, ref_info = _run(mesh, s0, dt, nsteps, args.reference, True, args.chunk, args.reference)
    f32, f32_info = _run(mesh, s0, dt, nsteps, "fp32", True, args.chunk, "fp32")
    out = {"n": args.n, "days": args.days, "steps": nsteps, "dt_s": dt, "reference": args.reference,
           "fp32_vs_fp64": _norms(mesh, f32, ref), "fp32": f32_info, args.reference: ref_info}
    if not args.no_floor:
        rng = np.random.default_rng(2024)
        sp = s0.copy()
        for k in ("phiG", "u", "v"):
            a = getattr(sp, k)
            a *= 1.0 + 2.0 ** -24 * rng.standard_normal(a.shape)
        pert, pert_info = _run(mesh, sp, dt, nsteps, args.reference, False, args.chunk, "perturbed")
        out["perturbed_fp64_vs_fp64"] = _norms(mesh, pert, ref)
        out["perturbed"] = pert_info
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
