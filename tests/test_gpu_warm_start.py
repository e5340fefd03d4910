"""Learned CG warm start on the GPU: the BF16-MFMA forward and backward
kernels against the FP64 oracle (with and without the device's BF16
roundings), training on samples the solver records, and the step with the
network in charge of the start, judged by the FP64 true residual."""
import numpy as np
import pytest

import silk_physics_amd as sp
from oracle import warm_start as W
from silk_physics_amd import scenes

from helpers import gpu_world

pytestmark = pytest.mark.gpu


def _net(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=W.N_PARAMS) * scale).astype(np.float32)


def _samples(seed, n):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 16), np.float32)
    X[:, :8] = rng.uniform(-1, 1, (n, 8))
    t = rng.normal(size=n).astype(np.float32)
    return X, t


@pytest.mark.parametrize("seed,n", [(0, 16), (1, 1000), (2, 4099)])
def test_forward_backward_match_oracle(seed, n):
    w = gpu_world()
    p = _net(seed)
    assert w.set_warm_start_net(p)
    X, t = _samples(seed + 10, n)
    y, g, loss = w.debug_warm_net(X, t)
    # Against the oracle with the device's BF16 roundings: FP32 against FP64
    # accumulation (a rounding decided the other way moves one activation by
    # a BF16 ulp, so the bounds are a little above FP32 level).
    lo, go, yo = W.loss_and_grad(p.astype(np.float64), X, t, emulate=True)
    ey = np.abs(y - yo).max() / max(1.0, np.abs(yo).max())
    eg = np.abs(g - go).max() / np.abs(go).max()
    print(f"n={n}: |y - y_oracle| {ey:.2e}, |grad - grad_oracle| {eg:.2e}, "
          f"loss {loss:.6f} vs {lo:.6f}")
    assert ey <= 2e-3
    assert eg <= 5e-3
    assert abs(loss - lo) <= 2e-3 * lo
    # Against the plain FP64 network: BF16-level agreement.
    lf, gf, yf = W.loss_and_grad(p.astype(np.float64), X, t)
    assert np.abs(y - yf).max() / max(1.0, np.abs(yf).max()) <= 5e-2
    assert np.abs(g - gf).max() / np.abs(gf).max() <= 1e-1


def _sheet_world(n, kw=None):
    V, F = scenes.grid_sheet(n, z=0.35)
    SV, SF = scenes.icosphere(2, 0.3, (0.5, 0.5, 0.0))
    g = dict(acceleration_z=-9.81, dt=1 / 60, max_inner_iteration=10,
             linear_solver_rel_tol_min=1e-6, linear_solver_rel_tol_max=1e-6,
             initial_linear_rel_tol=1e-6, pcg_criterion=sp.PCG_TRUE)
    g.update(kw or {})
    w = gpu_world(**g)
    cc = sp.CollisionConfig()
    r, h = w.add_cloth(sp.ClothConfig(), cc, V, F, [])
    assert r, r.to_string()
    r, _ = w.add_obstacle(cc, SV, SF)
    assert r, r.to_string()
    return w, h, V, g


def test_record_and_train_reduce_the_loss():
    w, h, V, g = _sheet_world(24)
    assert w.warm_start_record(200_000, 0.25)
    for _ in range(12):
        assert w.solver_step()
    n = w.warm_start_samples()
    print(f"recorded samples: {n}")
    assert n > 10_000
    r, loss = w.warm_start_train(300, 4096, 2e-3, seed=1)
    assert r, r.to_string()
    first, last = float(loss[:10].mean()), float(loss[-10:].mean())
    print(f"training loss {first:.4f} -> {last:.4f}")
    assert np.isfinite(loss).all()
    assert last < 0.7 * first
    r, p = w.get_warm_start_net()
    assert r and np.isfinite(p).all()


def test_projected_starts_cut_iterations():
    """The same 20-step drape from the plain start (the previous iterate, as
    the reference), the extrapolated start (x0 + c2 * previous correction)
    and the learned start (x0 + c1 * network + c2 * previous correction),
    both projections: every solve still meets the true-residual test (FP64
    residual of the iterate <= 1e-6 |b|, restarts included), the judge keeps
    a projected start only where it lowers the FP64 residual, and the
    iterations per solve are reported for each."""
    w, h, V, g = _sheet_world(24)
    assert w.warm_start_record(200_000, 0.25)
    for _ in range(12):
        assert w.solver_step()
    r, _ = w.warm_start_train(400, 4096, 2e-3, seed=2)
    assert r
    r, net = w.get_warm_start_net()
    assert w.warm_start_record(0, 0.0)

    def run(mode, proj=sp.PROJECT_RESIDUAL):
        ww, hh, VV, gg = _sheet_world(24)
        if mode == sp.WARM_LEARNED:
            assert ww.set_warm_start_net(net)
        assert ww.set_global_config(sp.GlobalConfig(**gg, warm_start=mode,
                                                    warm_projection=proj))
        it = sol = kept = ev = 0
        res = 0.0
        for _ in range(20):
            assert ww.solver_step()
            st = ww.last_stats
            it += st.pcg_iterations
            sol += st.pcg_solves
            kept += st.warm_start_kept
            ev += st.warm_start_solves
            res = max(res, st.max_true_rel_residual)
            assert st.pcg_stalled_solves == 0
            assert st.pcg_unverified_solves == 0
        return it / sol, sol, kept, ev, res

    out = {}
    for name, mode, proj in [("previous iterate", sp.WARM_PREVIOUS, 0),
                             ("extrapolated, min residual", sp.WARM_EXTRAPOLATE, 0),
                             ("extrapolated, Galerkin", sp.WARM_EXTRAPOLATE, 1),
                             ("learned, min residual", sp.WARM_LEARNED, 0),
                             ("learned, Galerkin", sp.WARM_LEARNED, 1)]:
        ips, sol, kept, ev, res = run(mode, proj)
        out[name] = ips
        print(f"{name}: {ips:.2f} PCG iterations per solve over {sol} solves, "
              f"projected start kept in {kept} of {ev}, max true residual {res:.2e}")
        assert res <= 1e-6
        assert 0 <= kept <= ev
        if mode != sp.WARM_PREVIOUS:
            assert ev > 0


def test_c5_learned_start_saves_iterations():
    """C5: the network trained on samples recorded from a 96x96 and a 112x112
    drape (the solver's own solves), then the 128x128 colliding sheet run
    20 steps from the plain start and from the learned start: both meet the
    true-residual test on every solve; the iterations per solve and the PCG
    time per solve of each are printed (what the network gains on C5)."""
    w, h, V, g = _sheet_world(96)
    assert w.warm_start_record(400_000, 0.1)
    for _ in range(10):
        assert w.solver_step()
    n1 = w.warm_start_samples()
    r, loss = w.warm_start_train(600, 8192, 2e-3, seed=5)
    assert r
    r, net = w.get_warm_start_net()
    print(f"trained on {n1} samples: loss {loss[:20].mean():.3f} -> {loss[-20:].mean():.3f}")

    def run(mode):
        ww, hh, VV, gg = _sheet_world(128, dict(max_inner_iteration=10))
        if mode == sp.WARM_LEARNED:
            assert ww.set_warm_start_net(net)
        assert ww.set_global_config(sp.GlobalConfig(**gg, warm_start=mode))
        ww.set_timing(True)
        ww.timing(1, reset=True)
        it = sol = kept = ev = 0
        for _ in range(20):
            assert ww.solver_step()
            st = ww.last_stats
            it += st.pcg_iterations
            sol += st.pcg_solves
            kept += st.warm_start_kept
            ev += st.warm_start_solves
            assert st.pcg_unverified_solves == 0 and st.pcg_stalled_solves == 0
            assert st.max_true_rel_residual <= 1e-6
        ms, n = ww.timing(1)
        return it / sol, ms / max(1, sol), kept, ev

    res = {}
    for name, mode in [("plain", sp.WARM_PREVIOUS), ("extrapolated", sp.WARM_EXTRAPOLATE),
                       ("learned", sp.WARM_LEARNED)]:
        ips, mps, kept, ev = run(mode)
        res[name] = ips
        print(f"C5 {name}: {ips:.2f} PCG iterations per solve, {1e3 * mps:.1f} us of PCG "
              f"per solve, projected start kept in {kept} of {ev}")
    print(f"learned / plain iterations: {res['learned'] / res['plain']:.3f}; "
          f"learned / extrapolated: {res['learned'] / res['extrapolated']:.3f}")
