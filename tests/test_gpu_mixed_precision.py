"""Mixed-precision MAS on the 256x256 sheet (66 049 vertices, seeded +-1 mm
jitter): the bank inverses stored in FP16 (scaled per bank) or BF16, products
accumulated in FP32, dot products in FP64 (the reference's FP32 banks:
mas_preconditioner.cu:657-702,750-831).  The preconditioner's precision may
change the iteration count, never the answer: each form is held to the FP64
oracle's PCG solution of the same system, and the whole step to the FP32
form.  At this size the persistent PCG needs more workgroups than the
device has CUs: it runs the variant that puts two workgroups on a CU for
FP32 and FP16 banks; BF16 banks on that variant take the multi-launch PCG
(their persistent solve of this sheet diverged)."""
import numpy as np
import pytest

import oracle as O
import silk_physics_amd as sp
from silk_physics_amd import scenes

from helpers import f32, gpu_world, inside_convex

pytestmark = pytest.mark.gpu

N = 256
SEED = 0x5117C0DE
PRECISIONS = [("fp32", sp.MAS_FP32), ("fp16", sp.MAS_FP16), ("bf16", sp.MAS_BF16)]


def c3_mesh():
    V, F = scenes.grid_sheet(N, z=0.31)
    rng = np.random.default_rng(SEED)
    V = (V + rng.uniform(-1e-3, 1e-3, V.shape)).astype(np.float32)
    return V, F


def test_c3_pcg_each_precision_matches_fp64_oracle():
    V, F = c3_mesh()
    w = gpu_world(pcg_criterion=sp.PCG_RHS, linear_solver_abs_tol=1e-30)
    r, h = w.add_cloth(sp.ClothConfig(), sp.CollisionConfig(is_collision_on=False,
                                                            is_self_collision_on=False),
                       V, F, [])
    assert r
    info = w.scene_info()
    print(f"C3 scene: {info['pcg_workgroups']} persistent workgroups in "
          f"{info['pcg_launches']} launch(es) (0: multi-launch PCG)")
    r, perm, offs = w.get_partition(h, V.shape[0])
    nv = V.shape[0]
    Vp = np.zeros_like(V)
    Vp[perm] = V
    Fp = perm[F]
    mass = 0.1 * O.mass_area(Vp.astype(np.float64), Fp)[0]
    d = (mass / f32(1 / 60) ** 2).astype(np.float32)
    rng = np.random.default_rng(7)
    b = np.zeros((nv, 3))
    b[:, 2] = -9.81 * mass
    b += rng.normal(size=(nv, 3)) * 1e-5
    b = b.astype(np.float32).reshape(-1)
    xo, oi = O.pcg(Vp.astype(np.float64), Fp, f32(100.0), f32(1e-4), offs,
                   np.repeat(d.astype(np.float64), 3), b.astype(np.float64),
                   np.zeros(b.size), 1e-6, 1e-30, 3000, criterion=1)
    assert oi["status"] == 1
    scale = np.abs(xo).max()
    print(f"FP64 oracle: {oi['iters']} iterations, max |x| {scale:.3e}")
    errs = {}
    for name, prec in PRECISIONS:
        assert w.set_global_config(sp.GlobalConfig(
            pcg_criterion=sp.PCG_RHS, linear_solver_abs_tol=1e-30, mas_precision=prec))
        x, it = w.debug_pcg(d, b, 1e-6)
        e = np.abs(x.astype(np.float64) - xo).reshape(-1, 3).max(axis=1)
        rms = float(np.sqrt(np.mean(e ** 2)))
        errs[name] = e.max()
        print(f"{name} banks: {it} iterations; per-vertex error vs FP64 oracle "
              f"max {e.max():.3e} RMS {rms:.3e} (relative {e.max() / scale:.2e})")
        assert it < 3000
        assert e.max() <= 2e-3 * scale
    # The bank precision moves the answer by no more than the tolerance does.
    assert errs["fp16"] <= 2 * errs["fp32"] + 1e-4 * scale
    assert errs["bf16"] <= 2 * errs["fp32"] + 1e-4 * scale


def test_c3_steps_each_precision():
    """Three steps of the C3 drape (sheet 1 cm above the sphere's top) per
    bank precision under the true-residual criterion: every solve's iterate
    meets ||b - A (x0 + dx)|| <= 1e-6 ||b|| (FP64; the FP32 recurrence alone
    stops at about 1.6e-6 here), no vertex enters the sphere, and the
    positions agree with the FP32-bank form."""
    V, F = c3_mesh()
    SV, SF = scenes.icosphere(4, 0.3, (0.5, 0.5, 0.0))
    out = {}
    for name, prec in PRECISIONS:
        w = gpu_world(acceleration_z=-9.81, dt=1 / 60, max_inner_iteration=10,
                      linear_solver_rel_tol_min=1e-6, linear_solver_rel_tol_max=1e-6,
                      initial_linear_rel_tol=1e-6, pcg_criterion=sp.PCG_TRUE,
                      mas_precision=prec, true_residual_stats=1)
        cc = sp.CollisionConfig()
        r, h = w.add_cloth(sp.ClothConfig(), cc, V, F, [])
        assert r
        r, _ = w.add_obstacle(cc, SV, SF)
        assert r
        its = sols = rst = 0
        res = 0.0
        for _ in range(3):
            assert w.solver_step()
            st = w.last_stats
            its += st.pcg_iterations
            sols += st.pcg_solves
            rst += st.pcg_restarts
            res = max(res, st.max_true_rel_residual)
            assert st.pcg_stalled_solves == 0
            assert st.pcg_unverified_solves == 0
        _, x = w.get_cloth_position(h, n=V.size)
        dmin = inside_convex(x, SV, SF).min()
        print(f"{name} banks: {its / max(1, sols):.1f} PCG iterations per solve over "
              f"{sols} solves ({rst} restarts), max true residual {res:.2e}, closest "
              f"vertex {dmin:+.2e} m outside the sphere")
        assert res <= 1e-6
        assert dmin > 0.0
        out[name] = x.astype(np.float64)
    for name in ("fp16", "bf16"):
        diff = np.abs(out[name] - out["fp32"]).max()
        print(f"{name} vs fp32 banks after 3 steps: max |dx| {diff:.2e} m")
        assert diff <= 1e-4
