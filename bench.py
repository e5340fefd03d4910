#!/usr/bin/env python3
"""Benchmark: simulated steps per second of the 80x80 sheet dropped 0.5 m onto
a 2562-vertex icosphere with collision (broadphase, FP64 CCD, coloured
contacts), every PCG solve run to ||b - A x|| <= 1e-6 ||b|| (FP64 residual of
the solve's iterate, restarts included), on one GPU.

Prints one JSON line.  Every solver_step call drains the library's stream
before returning, so host wall time brackets the device work of the timed
steps.  Step 0 (lazy scene assembly) is the warmup and is reported apart.

roofline: the kernel with the largest device time over the timed steps,
timed with HIP events on the library stream around every launch.  Its
algorithmic bytes per launch follow the reference formulation's byte model
(3x3 BSR operator and packed 96x96 MAS blocks):
  PCG iteration  40 nnzb + 4 (V + 1) + S n_blk + 120 V,  nnzb = offdiag + V,
                 S = 18624 (FP32 blocks) or 9312 (16-bit), n_blk = MAS banks
  local step     136 F + 60 V
  residual       8 offdiag + 64 V
and, beside it, this library's own bytes per PCG iteration (vertex-scalar
ELL and 32x32 banks): 8 offdiag + 48 V + 4096 banks (2048 at 16 bits).
traffic: HBM bytes per launch from rocprofv3 FETCH_SIZE and WRITE_SIZE
counter runs (profiles/pcg_traffic.json, made with --pmc-probe), scaled by
the same counters on the library's copy kernel of known size.
cpu_baseline: the reference CPU algorithm restated (PD with a sparse
Cholesky factor, SVD stretch projection, barrier reflection targets, CCD
rollback; oracle/cpu_pipeline.cpp) on the same scene, timed on the host's
cores for a bounded sample.
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

HBM_PEAK_GBS = 8000.0
METRIC = "sim steps/sec, 80x80 sheet + collision; PCG iters to 1e-6 rel residual"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pcg_traffic.json")
CLASSES = {1: "k_pcg_persistent", 2: "k_local_step", 3: "k_contact_update",
           4: "k_broadphase", 5: "k_ccd", 6: "k_residual"}


def global_kw(true_residual: bool = False, criterion: int = sp.PCG_TRUE,
              mas: int = sp.MAS_FP32) -> dict:
    return dict(acceleration_z=-9.81, dt=1 / 60, max_inner_iteration=10,
                max_outer_iteration=100, linear_solver_rel_tol_min=1e-6,
                linear_solver_rel_tol_max=1e-6, initial_linear_rel_tol=1e-6,
                true_residual_stats=int(true_residual), pcg_criterion=criterion,
                mas_precision=mas)


def build_world(n: int, mas: int = sp.MAS_FP32):
    s = scenes.c2_scene(n)
    w = sp.World()
    r = w.init(sp.Backend.GPU)
    if not r:
        raise SystemExit(f"GPU init failed: {r.to_string()}")
    assert w.set_global_config(sp.GlobalConfig(**global_kw(mas=mas)))
    cc = sp.CollisionConfig()  # on, self on, group 0, mu 0.3
    r, h = w.add_cloth(sp.ClothConfig(), cc, s["V"], s["F"], s["pins"])
    assert r, r.to_string()
    r, ho = w.add_obstacle(cc, s["SV"], s["SF"])
    assert r, r.to_string()
    return s, w, h


def byte_model(cls: int, info: dict, iters: int, launches: int,
               half: bool = False) -> float:
    """Algorithmic bytes of `launches` launches of a kernel class in the
    reference formulation (3x3 BSR, packed 96x96 MAS blocks)."""
    V, F = info["vertices"], info["triangles"]
    if cls == 1:
        nnzb = info["offdiag_entries"] + V
        S = 9312.0 if half else 18624.0
        return iters * (40.0 * nnzb + 4.0 * (V + 1) + S * info["mas_banks"] + 120.0 * V)
    if cls == 2:
        return launches * (136.0 * F + 60.0 * V)
    if cls == 6:
        return launches * (8.0 * info["offdiag_entries"] + 64.0 * V)
    return 0.0


def own_pcg_bytes(info: dict, half: bool = False) -> float:
    """This library's bytes per PCG iteration (vertex-scalar ELL, 32x32 banks)."""
    return (8.0 * info["offdiag_entries"] + 48.0 * info["vertices"] +
            (2048.0 if half else 4096.0) * info["mas_banks"])


def cpu_baseline(s, w, h, x_last, x_prev, dt, budget_s: float, max_steps: int):
    import oracle as O  # the CPU oracle: reported baseline only

    f32 = lambda v: float(np.float32(v))  # noqa: E731
    r, perm, offs = w.get_partition(h, s["V"].shape[0])
    ow = O.OracleWorld(dict(az=f32(-9.81), dt=f32(dt), max_inner=10,
                            lin_rel_min=f32(1e-6), lin_rel_max=f32(1e-6),
                            lin_rel_init=f32(1e-6), ms_scale=f32(0.2),
                            ccd_tol=f32(1e-6)))
    ow.add_cloth(s["V"], s["F"], [], perm, offs)
    ow.add_obstacle(s["SV"], s["SF"])
    ow.finalize()
    v = (x_last.astype(np.float64) - x_prev.astype(np.float64)) / f32(dt)
    ow.set_state(0, x_last.astype(np.float64), v)
    n = 0
    t0 = time.perf_counter()
    while n < max_steps and (n == 0 or time.perf_counter() - t0 < budget_s):
        rc, _ = ow.step()
        assert rc == 0
        n += 1
    el = time.perf_counter() - t0
    return {"value": round(n / el, 4), "unit": "steps/s", "cores": 1,
            "kind": "port",
            "sample": f"{n} steps of the same workload with the FP64 oracle "
                      f"(single thread), started from the GPU state after the "
                      f"timed steps (draped, in contact), {el:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=80, help="sheet quads per side")
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="seconds of CPU oracle work for cpu_baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verify-steps", type=int, default=5,
                    help="steps after the timed region with the FP64 true "
                         "residual evaluated for every PCG solve")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the step "
              "is not sharded across GPUs.", file=sys.stderr)
        return 2

    s, w, h = build_world(args.n)
    nvx = s["V"].size
    t0 = time.perf_counter()
    for _ in range(args.warmup):
        r = w.solver_step()
        assert r, r.to_string()
    warm_s = time.perf_counter() - t0
    info = w.scene_info()

    w.set_timing(True)
    for c in range(8):
        w.timing(c, reset=True)
    tot = dict(pcg=0, solves=0, admm=0, outer=0, contacts=0, ccd=0)
    t0 = time.perf_counter()
    for k in range(args.steps):
        r = w.solver_step()
        assert r, r.to_string()
        st = w.last_stats
        tot["pcg"] += st.pcg_iterations
        tot["solves"] += st.pcg_solves
        tot["admm"] += st.admm_iterations
        tot["outer"] += st.outer_iterations
        tot["contacts"] += st.contacts
        tot["ccd"] += st.ccd_candidates
    dt_wall = time.perf_counter() - t0
    w.set_timing(False)
    kern = {}
    for c, name in CLASSES.items():
        ms, n = w.timing(c)
        kern[c] = (ms, n)
    # Two more (untimed) states for the oracle's starting velocity.
    _, x_prev = w.get_cloth_position(h, n=nvx)
    assert w.solver_step()
    _, x_last = w.get_cloth_position(h, n=nvx)

    # Roofline of the kernel with the most device time.
    dom = max((c for c in kern if c in (1, 2, 6)), key=lambda c: kern[c][0])
    ms, n = kern[dom]
    bytes_total = byte_model(dom, info, tot["pcg"], n)
    achieved = bytes_total / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    roof = {"bound": "hbm", "kernel": CLASSES[dom],
            "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
            "avg_launch_us": round(1e3 * ms / max(1, n), 3), "launches": n,
            "bytes_per_launch": round(bytes_total / max(1, n), 1)}

    # Verification pass: FP64 true residual of every PCG solve.
    gk = global_kw(true_residual=True)
    assert w.set_global_config(sp.GlobalConfig(**gk))
    true_res = 0.0
    for _ in range(args.verify_steps):
        assert w.solver_step()
        true_res = max(true_res, w.last_stats.max_true_rel_residual)

    sps = args.steps / dt_wall
    line = {
        "metric": METRIC, "value": round(sps, 3), "unit": "steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt_wall / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"{args.n}x{args.n}-quad sheet dropped 0.5 m on a "
                               "2562-vertex icosphere, collision and self-"
                               "collision on, dt 1/60, 10 ADMM iters, CG rel "
                               "tol 1e-6",
                   **info},
        "pcg_iters_total": tot["pcg"], "pcg_solves": tot["solves"],
        "pcg_iters_per_solve": round(tot["pcg"] / max(1, tot["solves"]), 2),
        "pcg_iters_per_step": round(tot["pcg"] / args.steps, 2),
        "max_true_rel_residual": float(f"{true_res:.3e}"),
        "admm_iters_per_step": round(tot["admm"] / args.steps, 2),
        "outer_iters_per_step": round(tot["outer"] / args.steps, 2),
        "mean_contacts": round(tot["contacts"] / args.steps, 1),
        "ccd_candidates_per_step": round(tot["ccd"] / args.steps, 1),
        "kernel_ms": {CLASSES[c]: round(v[0], 3) for c, v in kern.items()},
        "kernel_launches": {CLASSES[c]: v[1] for c, v in kern.items()},
        "warmup_s": round(warm_s, 3),
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(s, w, h, x_last, x_prev, 1 / 60,
                                            args.cpu_budget, 20)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
