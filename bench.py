#!/usr/bin/env python3
"""Benchmark: simulated steps per second of the 80x80 sheet dropped 0.5 m onto
a 2562-vertex icosphere with collision (broadphase, FP64 CCD, coloured
contacts), every PCG solve run to ||b - A x|| <= 1e-6 ||b|| (FP64 residual of
the solve's iterate, restarts included), on one GPU.

Prints one JSON line.  Every solver_step call drains the library's stream
before returning, so host wall time brackets the device work of the timed
steps.  Step 0 (lazy scene assembly) is the warmup and is reported apart,
together with the launch autotuning that follows it (broadphase split depth
and PCG variant, which leave the results unchanged; --no-autotune skips it).

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
counter runs (profiles/k_pcg_persistent_c2_traffic.json, made with --pmc-probe), scaled by
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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_pcg_persistent_c2_traffic.json")
CLASSES = {1: "k_pcg_persistent", 2: "k_local_step", 3: "k_contact_update",
           4: "k_broadphase", 5: "k_ccd", 6: "k_residual"}


def global_kw(true_residual: bool = False, criterion: int = sp.PCG_TRUE,
              mas: int = sp.MAS_FP32) -> dict:
    return dict(acceleration_z=-9.81, dt=1 / 60, max_inner_iteration=10,
                max_outer_iteration=100, linear_solver_rel_tol_min=1e-6,
                linear_solver_rel_tol_max=1e-6, initial_linear_rel_tol=1e-6,
                true_residual_stats=int(true_residual), pcg_criterion=criterion,
                mas_precision=mas)


def build_world(n: int, mas: int = sp.MAS_FP32, batched: int = 0, jitter_mm: float = 0.0):
    """The C2-style drop of an n x n sheet (its vertices moved by a seeded
    uniform +-jitter_mm in every coordinate), or with batched > 0 that many
    independent sheets of 16..128 quads (each over its own sphere, in its own
    collision group).  Returns (scene of the first sheet, world, handle)."""
    w = sp.World()
    r = w.init(sp.Backend.GPU)
    if not r:
        raise SystemExit(f"GPU init failed: {r.to_string()}")
    assert w.set_global_config(sp.GlobalConfig(**global_kw(mas=mas)))
    sheets = scenes.c4_scene(batched) if batched else [scenes.c2_scene(n)]
    if jitter_mm > 0 and not batched:
        rng = np.random.default_rng(0x5117)
        V = sheets[0]["V"]
        sheets[0]["V"] = (V + rng.uniform(-1e-3 * jitter_mm, 1e-3 * jitter_mm,
                                          size=V.shape)).astype(np.float32)
    handles = []
    for s in sheets:
        cc = sp.CollisionConfig(group=s.get("group", 0))  # on, self on, mu 0.3
        r, h = w.add_cloth(sp.ClothConfig(), cc, s["V"], s["F"], s["pins"])
        assert r, r.to_string()
        r, ho = w.add_obstacle(cc, s["SV"], s["SF"])
        assert r, r.to_string()
        handles.append(h)
    return sheets[0], w, handles[0]


def train_warm_net(n: int = 96, steps: int = 10) -> np.ndarray:
    """The per-vertex warm-start network trained on the GPU from the solves
    of an n x n drop on the same sphere (samples recorded by the solver)."""
    s = scenes.c2_scene(n, z=0.32)
    w = sp.World()
    assert w.init(sp.Backend.GPU)
    assert w.set_global_config(sp.GlobalConfig(**global_kw()))
    cc = sp.CollisionConfig()
    r, h = w.add_cloth(sp.ClothConfig(), cc, s["V"], s["F"], s["pins"])
    assert r, r.to_string()
    r, _ = w.add_obstacle(cc, s["SV"], s["SF"])
    assert r, r.to_string()
    assert w.warm_start_record(400_000, 0.1)
    for _ in range(steps):
        assert w.solver_step()
    r, _ = w.warm_start_train(600, 8192, 2e-3, seed=5)
    assert r, r.to_string()
    r, net = w.get_warm_start_net()
    assert r
    return net


def compare_c5(args) -> int:
    """The 128x128 colliding sheet (reference ADMM defaults: up to 100
    iterations with the convergence test) in two forms at the same FP64 test
    of every PCG solve (||b - A x|| <= 1e-6 ||b||): the plain form (FP32 MAS,
    CG from the previous iterate) and the learned form (BF16 MAS, projected
    learned start).  One JSON line with each form's steps/s and PCG
    iterations per solve over the same steps from rest."""
    net = train_warm_net()
    out = {}
    for name, mas, warm in (("plain", sp.MAS_FP32, sp.WARM_PREVIOUS),
                            ("learned_bf16", sp.MAS_BF16, sp.WARM_LEARNED)):
        s = scenes.c2_scene(128, z=0.32)
        w = sp.World()
        assert w.init(sp.Backend.GPU)
        kw = global_kw(mas=mas)
        kw.pop("max_inner_iteration")
        if warm == sp.WARM_LEARNED:
            assert w.set_warm_start_net(net)
        assert w.set_global_config(sp.GlobalConfig(**kw, warm_start=warm))
        cc = sp.CollisionConfig()
        r, h = w.add_cloth(sp.ClothConfig(), cc, s["V"], s["F"], s["pins"])
        assert r, r.to_string()
        r, _ = w.add_obstacle(cc, s["SV"], s["SF"])
        assert r, r.to_string()
        assert w.solver_step()  # scene assembly
        it = sol = unver = 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            assert w.solver_step()
            st = w.last_stats
            it += st.pcg_iterations
            sol += st.pcg_solves
            unver += st.pcg_unverified_solves
        dt = time.perf_counter() - t0
        out[name] = {"steps_per_s": round(args.steps / dt, 3),
                     "pcg_iters_per_solve": round(it / max(1, sol), 2),
                     "pcg_solves": sol, "unverified": unver}
    out["ratio_steps_per_s"] = round(out["learned_bf16"]["steps_per_s"] /
                                     out["plain"]["steps_per_s"], 3)
    print(json.dumps({"comparison": "128x128 colliding sheet, learned + BF16 vs plain",
                      "steps": args.steps, **out}), flush=True)
    return 0


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


def host_cpu() -> dict:
    """Host core count and model name (from os and /proc/cpuinfo)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return {"nproc": cores, "model": model}


def cpu_baseline(n: int, budget_s: float, gpu_step_s: list) -> dict:
    """The reference CPU algorithm restated (oracle/cpu_pipeline.cpp) on the
    same scene from its rest state: steps/s over a bounded sample on all host
    cores, and on one thread; the GPU's rate over the same first steps is
    given beside it."""
    import oracle as O  # the restated reference CPU path: baseline only

    hc = host_cpu()
    out = {}
    for label, threads in (("all", hc["nproc"]), ("one", 1)):
        s = scenes.c2_scene(n)
        p = O.CpuPipeline(threads=threads)
        p.add_cloth(s["V"], s["F"], s["pins"])
        p.add_obstacle(s["SV"], s["SF"])
        k = 0
        t0 = time.perf_counter()
        while k < len(gpu_step_s) and (k == 0 or time.perf_counter() - t0 < budget_s):
            rc, _ = p.step()
            if rc != 0:
                raise RuntimeError("CPU pipeline step failed")
            k += 1
        out[label] = (k, time.perf_counter() - t0)
    k, el = out["all"]
    k1, el1 = out["one"]
    gpu_same = k / max(1e-9, sum(gpu_step_s[:k]))
    # The library's own CPU backend (Backend.CPU) on the same first k steps.
    s = scenes.c2_scene(n)
    wc = sp.World()
    assert wc.init(sp.Backend.CPU)
    assert wc.set_global_config(sp.GlobalConfig(acceleration_z=-9.81, dt=1 / 60,
                                                max_inner_iteration=10))
    cc = sp.CollisionConfig()
    assert wc.add_cloth(sp.ClothConfig(), cc, s["V"], s["F"], s["pins"])[0]
    assert wc.add_obstacle(cc, s["SV"], s["SF"])[0]
    t0 = time.perf_counter()
    for _ in range(k):
        assert wc.solver_step()
    el_lib = time.perf_counter() - t0
    return {"value": round(k / el, 4), "unit": "steps/s", "cores": hc["nproc"],
            "kind": "port", "one_thread": round(k1 / el1, 4),
            "library_cpu_backend": round(k / el_lib, 4),
            "host": hc["model"],
            "sample": f"steps 1..{k} of the same 80x80 drop from rest with the "
                      f"reference CPU algorithm restated (PD, envelope Cholesky, "
                      f"SVD projection, barrier targets, FP32 CCD) on "
                      f"{hc['nproc']} threads, {el:.1f} s; one thread: {k1} steps "
                      f"in {el1:.1f} s; the library's CPU backend (Backend.CPU, "
                      f"its own threads) ran them in {el_lib:.1f} s; the GPU ran "
                      f"the same first {k} steps at {gpu_same:.2f} steps/s"}


def traffic_summary() -> dict | None:
    """HBM bytes per launch of the persistent PCG from the committed counter
    summary (profiles/k_pcg_persistent_c2_traffic.json), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def pmc_probe(args) -> int:
    """Short run for the hardware-counter passes: C2 steps (PCG solves from
    rest, then the first contacts), reporting the PCG iterations and launches
    they took, then the library's copy kernel on 256 MiB three times as the
    reference of known traffic."""
    s, w, h = build_world(args.n)
    iters = 0
    for _ in range(args.steps):
        assert w.solver_step()
        iters += w.last_stats.pcg_iterations
    info = w.scene_info()
    ms = w.debug_copy(256 << 20, 3)
    print(json.dumps({"probe": "pmc", "steps": args.steps, "pcg_iterations": iters,
                      "copy_bytes": 256 << 20, "copy_reps": 3, "copy_ms": round(ms, 3),
                      **info}), flush=True)
    return 0


def pmc_summary(fetch_dir: str, write_dir: str, probe_log: str) -> int:
    """profiles/k_pcg_persistent_c2_traffic.json from two rocprofv3 --pmc runs of the probe
    (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch): the persistent PCG's
    bytes per iteration, each counter scaled by what it reads for the copy
    kernel's known 256 MiB per launch."""
    import csv
    import glob

    def per_kernel(d: str, counter: str) -> dict:
        out: dict = {}
        for fn in glob.glob(os.path.join(d, "**", "*counter_collection*.csv"), recursive=True):
            with open(fn) as f:
                for row in csv.DictReader(f):
                    if row.get("Counter_Name") != counter:
                        continue
                    name = row.get("Kernel_Name", "")
                    key = ("pcg" if "k_pcg_persistent" in name else
                           "copy" if "k_copy_unless" in name else None)
                    if key is None:
                        continue
                    disp = (fn, int(row.get("Dispatch_Id", 0)))
                    out.setdefault(key, {}).setdefault(disp, 0.0)
                    out[key][disp] += float(row["Counter_Value"])
        # In dispatch order; the reference copies are the probe's last ones
        # (the step also runs the copy kernel on small buffers).
        return {k: [v[d] for d in sorted(v)] for k, v in out.items()}

    probe = None
    with open(probe_log) as f:
        for line in f:
            if line.startswith("{") and '"probe"' in line:
                probe = json.loads(line)
    fetch = per_kernel(fetch_dir, "FETCH_SIZE")
    write = per_kernel(write_dir, "WRITE_SIZE")
    if probe is None or not fetch.get("pcg") or not fetch.get("copy") or \
            not write.get("pcg") or not write.get("copy"):
        print("pmc summary: missing probe line or counter rows", file=sys.stderr)
        return 1
    copy_kib = probe["copy_bytes"] / 1024.0
    reps = probe["copy_reps"]
    fetch["copy"], write["copy"] = fetch["copy"][-reps:], write["copy"][-reps:]
    f_scale = copy_kib / float(np.mean(fetch["copy"]))
    w_scale = copy_kib / float(np.mean(write["copy"]))
    pcg_fetch = float(np.sum(fetch["pcg"])) * f_scale * 1024.0
    pcg_write = float(np.sum(write["pcg"])) * w_scale * 1024.0
    iters = probe["pcg_iterations"]
    summary = {
        "kernel": "k_pcg_persistent", "vertices": probe["vertices"],
        "probe_steps": probe["steps"], "pcg_iterations": iters,
        "pcg_launches": len(fetch["pcg"]),
        "fetch_kib_per_launch_raw": round(float(np.mean(fetch["pcg"])), 1),
        "write_kib_per_launch_raw": round(float(np.mean(write["pcg"])), 1),
        "copy_fetch_kib_raw": round(float(np.mean(fetch["copy"])), 1),
        "copy_write_kib_raw": round(float(np.mean(write["copy"])), 1),
        "fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4),
        "bytes_per_iteration": round((pcg_fetch + pcg_write) / max(1, iters), 1),
        "fetch_bytes_per_iteration": round(pcg_fetch / max(1, iters), 1),
        "write_bytes_per_iteration": round(pcg_write / max(1, iters), 1),
        "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of bench.py --pmc-probe, "
                  "scaled by the copy kernel's counters for 256 MiB",
    }
    os.makedirs(os.path.dirname(TRAFFIC_FILE), exist_ok=True)
    with open(TRAFFIC_FILE, "w") as f:
        json.dump(summary, f, indent=1)
    print(json.dumps(summary))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=80, help="sheet quads per side")
    ap.add_argument("--mas", choices=["fp32", "fp16", "bf16"], default="fp32",
                    help="storage of the MAS bank inverses")
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="seconds of CPU work per thread count for cpu_baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verify-steps", type=int, default=5,
                    help="steps after the timed region with the FP64 residual of "
                         "the stored x reported, then with the reference's "
                         "stopping rule for its iteration count")
    ap.add_argument("--jitter", type=float, default=0.0,
                    help="seeded uniform vertex jitter of the sheet in mm (the "
                         "mixed-precision workload uses 1 mm at n = 256)")
    ap.add_argument("--batched", type=int, default=0,
                    help="instead of one sheet, this many independent sheets of "
                         "16..128 quads (the batched-instance scene)")
    ap.add_argument("--breakdown-steps", type=int, default=20,
                    help="steps after the timed region with every kernel class "
                         "timed by events (the timed region times only the PCG "
                         "solve kernel: events around every launch cost ~5 %%)")
    ap.add_argument("--graph", action="store_true",
                    help="capture each outer iteration's ADMM loop into a hipGraph "
                         "(admm_graph); the timed region then records no kernel "
                         "events, and the roofline takes the PCG kernel's time "
                         "from the breakdown pass")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes (the broadphase split "
                         "depth and the PCG variant are otherwise timed after "
                         "the warmup and the fastest kept)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="short run for rocprofv3 --pmc passes (no JSON metric)")
    ap.add_argument("--compare-c5", action="store_true",
                    help="the 128x128 sheet in the plain and the learned + BF16 "
                         "forms (one JSON line, not the metric)")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "PROBE_LOG"),
                    help="write profiles/k_pcg_persistent_c2_traffic.json from the two counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the step "
              "is not sharded across GPUs.",
              file=sys.stderr)
        return 2
    if args.pmc_probe:
        return pmc_probe(args)
    if args.pmc_summary:
        return pmc_summary(*args.pmc_summary)
    if args.compare_c5:
        return compare_c5(args)
    mas = {"fp32": sp.MAS_FP32, "fp16": sp.MAS_FP16, "bf16": sp.MAS_BF16}[args.mas]

    s, w, h = build_world(args.n, mas, args.batched, args.jitter)
    if args.batched:
        args.no_cpu_baseline = True
    t0 = time.perf_counter()
    for _ in range(args.warmup):
        r = w.solver_step()
        assert r, r.to_string()
    tune = None
    if not args.no_autotune:
        r, tune = w.autotune()
        assert r, r.to_string()
    warm_s = time.perf_counter() - t0
    info = w.scene_info()

    if args.graph:
        assert w.set_global_config(sp.GlobalConfig(**global_kw(False, sp.PCG_TRUE, mas),
                                                   admm_graph=sp.GRAPH_ON))
    # Only the PCG solve kernel (class 1) is timed over the timed region, for
    # the roofline; with the ADMM loop in a graph no class inside it is.
    w.set_timing(0 if args.graph else 1 << 1)
    for c in range(9):
        w.timing(c, reset=True)
    graph_launches = 0
    tot = dict(pcg=0, solves=0, restarts=0, admm=0, outer=0, contacts=0, ccd=0,
               unverified=0, stalled=0)
    step_s = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        r = w.solver_step()
        step_s.append(time.perf_counter() - ts)
        assert r, r.to_string()
        st = w.last_stats
        tot["pcg"] += st.pcg_iterations
        tot["solves"] += st.pcg_solves
        tot["restarts"] += st.pcg_restarts
        tot["admm"] += st.admm_iterations
        tot["outer"] += st.outer_iterations
        tot["contacts"] += st.contacts
        tot["ccd"] += st.ccd_candidates
        tot["unverified"] += st.pcg_unverified_solves
        tot["stalled"] += st.pcg_stalled_solves
        graph_launches += st.graph_launches
    dt_wall = time.perf_counter() - t0
    w.set_timing(False)
    pcg_timed = w.timing(1)
    if args.graph:
        assert w.set_global_config(sp.GlobalConfig(**global_kw(False, sp.PCG_TRUE, mas)))

    # Per-class device time over a short pass after the timed region, every
    # class timed by events around its launches.
    w.set_timing(True)
    for c in range(9):
        w.timing(c, reset=True)
    bd_pcg = 0
    for _ in range(args.breakdown_steps):
        assert w.solver_step()
        bd_pcg += w.last_stats.pcg_iterations
    w.set_timing(False)
    kern = {c: w.timing(c) for c in CLASSES}

    # Roofline of the PCG solve kernel, the one with the most device time
    # (breakdown pass), from its events over the timed region.
    half = mas != sp.MAS_FP32
    dom = max((c for c in kern if c in (1, 2, 6)), key=lambda c: kern[c][0])
    ms, n = pcg_timed
    pcg_iters = tot["pcg"]
    if dom != 1 or n == 0:
        ms, n = kern[dom]
        pcg_iters = bd_pcg
    # With several systems per launch the iteration total sums every
    # system's iterations, while the byte model covers all systems at once:
    # the model is applied to the iterations per system (an average over
    # systems of uneven size).
    nsys = max(1, info.get("systems", 1))
    pcg_iters_model = pcg_iters / nsys
    bytes_total = byte_model(dom, info, pcg_iters_model, n, half)
    achieved = bytes_total / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    roof = {"bound": "hbm", "kernel": CLASSES[dom],
            "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
            "avg_launch_us": round(1e3 * ms / max(1, n), 3), "launches": n,
            "bytes_per_launch": round(bytes_total / max(1, n), 1),
            "bytes_model": "reference formulation: 40 nnzb + 4 (V+1) + S n_blk + "
                           "120 V per PCG iteration"}
    roof["timed_over"] = ("the timed region" if pcg_iters == tot["pcg"] else
                          f"the {args.breakdown_steps}-step breakdown pass")
    if dom == 1 and ms > 0:
        own = own_pcg_bytes(info, half) * pcg_iters_model
        roof["own_bytes_per_launch"] = round(own / max(1, n), 1)
        roof["own_achieved"] = round(own / (ms * 1e-3) / 1e9, 2)
        roof["own_frac"] = round(own / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)
        roof["us_per_iteration"] = round(1e3 * ms / max(1e-9, pcg_iters_model), 3)
        tr = traffic_summary()
        if tr is not None and tr.get("vertices") == info["vertices"]:
            # Counter bytes per PCG iteration (probe run) times this run's
            # iterations per launch.
            roof["traffic"] = round(tr["bytes_per_iteration"] * pcg_iters_model / max(1, n), 1)
            roof["traffic_source"] = tr.get("source")

    # Verification: the FP64 residual of the stored FP32 x beside the
    # iterate's, then the reference's stopping rule (rr <= rel^2 rr0) for its
    # iteration count.
    assert w.set_global_config(sp.GlobalConfig(**global_kw(True, sp.PCG_TRUE, mas)))
    true_res = commit_res = 0.0
    for _ in range(args.verify_steps):
        assert w.solver_step()
        true_res = max(true_res, w.last_stats.max_true_rel_residual)
        commit_res = max(commit_res, w.last_stats.max_committed_rel_residual)
    assert w.set_global_config(sp.GlobalConfig(**global_kw(False, sp.PCG_RR0, mas)))
    ref_it = ref_sol = 0
    for _ in range(args.verify_steps):
        assert w.solver_step()
        ref_it += w.last_stats.pcg_iterations
        ref_sol += w.last_stats.pcg_solves

    sps = args.steps / dt_wall
    line = {
        "metric": METRIC, "value": round(sps, 3), "unit": "steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt_wall / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": (f"{args.batched} independent sheets of 16..128 quads "
                                "(seeded) 2 cm over their own 642-vertex spheres"
                                if args.batched else
                                f"{args.n}x{args.n}-quad sheet"
                                + (f" (+-{args.jitter:g} mm vertex jitter)" if args.jitter else "")
                                + " dropped 0.5 m on a 2562-vertex icosphere") +
                               ", collision and self-collision on, dt 1/60, 10 "
                               "ADMM iters, CG to ||b - A x|| <= 1e-6 ||b|| (FP64)",
                   "mas_precision": args.mas, **info},
        "pcg_iters_total": tot["pcg"], "pcg_solves": tot["solves"],
        "pcg_restarts": tot["restarts"],
        "pcg_iters_per_solve": round(tot["pcg"] / max(1, tot["solves"]), 2),
        "pcg_iters_per_step": round(tot["pcg"] / args.steps, 2),
        "pcg_unverified_solves": tot["unverified"],
        "pcg_stalled_solves": tot["stalled"],
        "max_true_rel_residual": float(f"{true_res:.3e}"),
        "max_stored_x_rel_residual": float(f"{commit_res:.3e}"),
        "reference_rule_iters_per_solve": round(ref_it / max(1, ref_sol), 2),
        "admm_iters_per_step": round(tot["admm"] / args.steps, 2),
        "outer_iters_per_step": round(tot["outer"] / args.steps, 2),
        "mean_contacts": round(tot["contacts"] / args.steps, 1),
        "ccd_candidates_per_step": round(tot["ccd"] / args.steps, 1),
        "admm_graph": bool(args.graph), "graph_launches": graph_launches,
        "breakdown_steps": args.breakdown_steps,
        "kernel_ms": {CLASSES[c]: round(v[0], 3) for c, v in kern.items()},
        "kernel_launches": {CLASSES[c]: v[1] for c, v in kern.items()},
        "warmup_s": round(warm_s, 3),
        "autotune": tune,
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.n, args.cpu_budget, step_s)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    # A successful run returns normally (the interpreter's own exit), so
    # that tools attached to the process finish their output first.
    rc = main()
    if rc:
        sys.exit(rc)
