"""Batched ensemble driver: many independent shallow-water runs at mixed
panel resolutions, each with its own seed (bumps and mountain) and its own
CFL time step, advanced together on one GPU.

``python -m conformal_cubed_sphere_amd.ensemble_run --members 64 --steps 60``
prints one JSON line: member cell-updates/s, the autotuner's choice and
per-member mass drift.  The reference runs one member per process
(ccsm.F90:2-73); each member here evolves exactly as that single run would.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from .core import GRAVITY, Ensemble, Mesh, cfl_time_step, init_case

SIZES = (48, 96, 144, 192)


def build_members(members: int, sizes=SIZES, size_seed: int = 7, days: float = 5.0, radius: float = 6.371e6,
                  omega: float = 7.292e-5):
    """Members of the mixed-resolution benchmark: N drawn uniformly from
    `sizes` (seeded), member m seeded with 1 + m, seeded balanced jet with 8
    bumps and a Gaussian mountain, dt from CFL 0.5 over `days`.
    Returns [(mesh, state, dt)] and the list of N."""
    rng = np.random.default_rng(size_seed)
    ns = [int(x) for x in rng.choice(list(sizes), size=members)]
    meshes = {n: Mesh(n, halo=3, radius=radius, omega=omega) for n in sorted(set(ns))}
    runs = []
    for m, n in enumerate(ns):
        s0 = init_case(meshes[n], case_num=7, seed=1 + m, n_bumps=8, mountain=True, gravity=GRAVITY)
        dt, _ = cfl_time_step(meshes[n], s0, 0.5, days * 86400.0)
        runs.append((meshes[n], s0, dt))
    return runs, ns


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--settle", type=float, default=0.3, help="seconds of untimed steps before the warmup")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64_fast"])
    ap.add_argument("--scheme", default="RK3_TVD", choices=["RK3_TVD", "RK4"])
    ap.add_argument("--tile-y", type=int, default=0, help="8, 16, 32 or 0 (autotune)")
    ap.add_argument("--persistent", type=int, default=-1, help="0, 1 or -1 (autotune)")
    ap.add_argument("--sizes", default="48,96,144,192", help="panel sizes the members draw from")
    args = ap.parse_args()
    runs, ns = build_members(args.members, sizes=tuple(int(x) for x in args.sizes.split(",")))
    cells = sum(6 * n * n for n in ns)
    with Ensemble([r[0] for r in runs], [r[2] for r in runs], precision=args.precision, scheme=args.scheme,
                  tile_y=args.tile_y, persistent=args.persistent) as ens:
        for k, r in enumerate(runs):
            ens.set_state(k, r[1])
        m0 = [ens.diagnostics(k)[0] for k in range(len(ens))]
        choice = ens.autotune()
        # untimed steps until the GPU has run for --settle seconds: its clocks
        # ramp up over the first tens of milliseconds of sustained work, and
        # drop again while it idles (re-uploading the states would idle it)
        settle = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < args.settle:
            ens.step(16)
            ens.synchronize()
            settle += 16
        ens.step(args.warmup)
        ens.synchronize()
        t0 = time.perf_counter()
        ens.step(args.steps)
        ens.synchronize()
        wall = time.perf_counter() - t0
        drift = [abs(ens.diagnostics(k)[0] - m0[k]) / m0[k] for k in range(len(ens))]
    print(json.dumps({"members": args.members, "interior_cells": cells, "steps": args.steps,
                      "member_cell_updates_per_s": cells * args.steps / wall, "ms_per_step": wall * 1e3 / args.steps,
                      "autotune": choice, "precision": args.precision, "scheme": args.scheme,
                      "mass_drift_max": max(drift), "mass_drift_after_steps": settle + args.warmup + args.steps}),
          flush=True)


if __name__ == "__main__":
    main()
