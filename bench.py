#!/usr/bin/env python3
"""Benchmark: simulated steps per second of the 80x80 sheet dropped onto a
sphere obstacle with collision (broadphase, FP64 CCD, coloured contacts),
PCG to 1e-6 relative residual, on one GPU.

Prints one JSON line (see --help).  Every solver_step call drains the
library's stream before returning, so host wall time brackets device work.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import silk_physics_amd as sp  # noqa: E402
from silk_physics_amd import scenes  # noqa: E402


def build_world(n: int, true_residual: bool = False):
    s = scenes.c2_scene(n)
    w = sp.World()
    r = w.init(sp.Backend.GPU)
    if not r:
        raise SystemExit(f"GPU init failed: {r.to_string()}")
    g = sp.GlobalConfig(acceleration_z=-9.81, dt=1 / 60, max_inner_iteration=10,
                        max_outer_iteration=100, linear_solver_rel_tol_min=1e-6,
                        linear_solver_rel_tol_max=1e-6, initial_linear_rel_tol=1e-6,
                        true_residual_stats=int(true_residual))
    assert w.set_global_config(g)
    cc = sp.CollisionConfig()  # on, self on, group 0, mu 0.3
    r, h = w.add_cloth(sp.ClothConfig(), cc, s["V"], s["F"], s["pins"])
    assert r, r.to_string()
    r, ho = w.add_obstacle(cc, s["SV"], s["SF"])
    assert r, r.to_string()
    return s, w, h


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=80, help="sheet quads per side")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the step "
              "is not sharded across GPUs.", file=sys.stderr)
        return 2

    s, w, h = build_world(args.n)
    t0 = time.perf_counter()
    for _ in range(args.warmup):
        r = w.solver_step()
        assert r, r.to_string()
    warm_s = time.perf_counter() - t0

    tot = dict(pcg=0, solves=0, admm=0, outer=0, contacts=0, ccd=0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = w.solver_step()
        assert r, r.to_string()
        st = w.last_stats
        tot["pcg"] += st.pcg_iterations
        tot["solves"] += st.pcg_solves
        tot["admm"] += st.admm_iterations
        tot["outer"] += st.outer_iterations
        tot["contacts"] += st.contacts
        tot["ccd"] += st.ccd_candidates
    dt = time.perf_counter() - t0
    sps = args.steps / dt
    line = {
        "metric": "sim steps/sec, 80x80 sheet + collision; PCG iters to 1e-6 rel residual",
        "value": round(sps, 3), "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{args.n}x{args.n}-quad sheet dropped 0.5 m on a "
                               "2562-vertex icosphere, collision on, dt 1/60, "
                               "10 ADMM iters, CG rel tol 1e-6",
                   "sheet_vertices": int(s["V"].shape[0]),
                   "sheet_triangles": int(s["F"].shape[0])},
        "pcg_iters_total": tot["pcg"], "pcg_solves": tot["solves"],
        "pcg_iters_per_solve": round(tot["pcg"] / max(1, tot["solves"]), 2),
        "admm_iters_per_step": round(tot["admm"] / args.steps, 2),
        "outer_iters_per_step": round(tot["outer"] / args.steps, 2),
        "mean_contacts": round(tot["contacts"] / args.steps, 1),
        "warmup_s": round(warm_s, 3),
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
