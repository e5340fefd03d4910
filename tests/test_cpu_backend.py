"""The library's CPU backend (World.init(Backend.CPU), silk::Backend::CPU):
the reference's CPU solver step (silk/src/backend/cpu/solver/pipeline.cpp)
on the host, through the same C ABI as the GPU backend.

Checked against the NumPy restatement of the reference CPU step
(oracle/cpu_pd.py) whose pieces are pinned first: its projection against the
oracle's project_stretch and against the closed form of an isotropic
stretch, its step against implicit-Euler free fall and a pinned sheet at
rest.  With contacts, the CPU backend is held to the restated reference CPU
pipeline (oracle CpuPipeline) before contact and to exact no-penetration
checks after it."""
import numpy as np
import pytest

import oracle as O
from oracle.cpu_pd import CpuPD, project_batch
import silk_physics_amd as sp
from silk_physics_amd import scenes

from helpers import inside_convex, self_intersections


def cpu_world(**kw):
    w = sp.World()
    r = w.init(sp.Backend.CPU)
    assert r, r.to_string()
    g = dict(acceleration_z=-9.81, dt=1 / 60)
    g.update(kw)
    r = w.set_global_config(sp.GlobalConfig(**g))
    assert r, r.to_string()
    return w


# ------------------------------------------------------------ oracle pins

def test_oracle_projection_matches_project_stretch():
    rng = np.random.default_rng(3)
    D = rng.normal(size=(200, 6))
    D[:50] *= 0.3   # compressed
    D[50:100] *= 2  # stretched
    T = project_batch(D)
    for d, t in zip(D, T):
        np.testing.assert_allclose(t, O.project_stretch(d), atol=1e-12)


def test_oracle_projection_closed_form():
    # F = s R (columns of a rotation scaled by s): T = clamp(s) R.
    rng = np.random.default_rng(4)
    for s in (0.5, 0.95, 1.0, 1.05, 1.7):
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        R = Q[:, :2]
        D = (s * R).T.reshape(-1)  # columns stored one after the other
        T = project_batch(D[None])[0]
        np.testing.assert_allclose(T, (np.clip(s, 0.9, 1.1) * R).T.reshape(-1), atol=1e-12)


def test_oracle_free_fall_closed_form():
    V, F = scenes.grid_sheet(6, z=0.2)
    o = CpuPD(V, F, damping=0.01)
    x0 = o.x.copy()
    v = np.zeros(3)
    x = np.zeros(3)
    h, g = o.h, np.array([0.0, 0.0, -9.81])
    for _ in range(6):
        o.step()
        v = 0.99 * v + h * g
        x = x + h * v
    np.testing.assert_allclose(o.x.reshape(-1, 3) - x0.reshape(-1, 3), np.tile(x, (V.shape[0], 1)),
                               atol=1e-9)
    assert o.last_inner == 1


def test_oracle_pinned_sheet_at_rest():
    V, F = scenes.grid_sheet(6)
    o = CpuPD(V, F, pins=scenes.top_corner_pins(6), g=(0.0, 0.0, 0.0))
    for _ in range(3):
        o.step()
    np.testing.assert_allclose(o.x, V.reshape(-1).astype(np.float64), atol=1e-9)


# ---------------------------------------------------- product vs oracle

def test_cpu_backend_free_fall_closed_form():
    V, F = scenes.grid_sheet(10, z=0.2)
    w = cpu_world()
    cc = sp.CollisionConfig(is_collision_on=False)
    r, h = w.add_cloth(sp.ClothConfig(), cc, V, F, [])
    assert r, r.to_string()
    v = 0.0
    z = 0.0
    for _ in range(10):
        assert w.solver_step()
        v = 0.99 * v - 9.81 / 60
        z += v / 60
    _, x = w.get_cloth_position(h, n=V.size)
    d = x.reshape(-1, 3) - V
    assert np.abs(d[:, :2]).max() < 1e-5
    assert np.abs(d[:, 2] - z).max() < 1e-5, (d[:, 2].min(), d[:, 2].max(), z)


@pytest.mark.parametrize("n,steps,max_inner", [(8, 60, 100), (32, 200, 10)])
def test_cpu_backend_matches_oracle_pinned_sheet(n, steps, max_inner):
    """C1 at n = 32: the 32x32-quad sheet, two top corners pinned, 200 steps
    of 1/60 s, every step within 1e-4 m of the FP64 restatement."""
    V, F = scenes.grid_sheet(n)
    pins = scenes.top_corner_pins(n)
    w = cpu_world(max_inner_iteration=max_inner)
    r, h = w.add_cloth(sp.ClothConfig(), sp.CollisionConfig(is_collision_on=False), V, F, pins)
    assert r, r.to_string()
    o = CpuPD(V, F, pins=pins, max_inner=max_inner)
    worst = 0.0
    for k in range(steps):
        assert w.solver_step(), k
        o.step()
        assert w.last_stats.admm_iterations == o.last_inner, k
        _, x = w.get_cloth_position(h, n=V.size)
        err = float(np.abs(x.astype(np.float64) - o.x).max())
        worst = max(worst, err)
        assert err <= 1e-4, (k, err)
    drop = float(V[:, 2].max() - x.reshape(-1, 3)[:, 2].min())
    print(f"{n}x{n}, {steps} steps: max |x_cpu - x_oracle| = {worst:.3e} m, "
          f"lowest vertex {drop:.3f} m below the rest plane")
    assert drop > 0.1  # the sheet swings down


def test_cpu_backend_matches_reference_pipeline_and_never_penetrates():
    """A 16x16 sheet dropped on an icosphere: identical to the restated
    reference CPU pipeline (FP32 CCD) while nothing is near the sphere, then
    no vertex inside the sphere and no self-intersection on any step."""
    V, F = scenes.grid_sheet(16, z=0.36)
    SV, SF = scenes.icosphere(2, 0.3, (0.5, 0.5, 0.0))
    w = cpu_world(max_inner_iteration=10)
    cc = sp.CollisionConfig()
    r, h = w.add_cloth(sp.ClothConfig(), cc, V, F, [])
    assert r, r.to_string()
    r, _ = w.add_obstacle(cc, SV, SF)
    assert r, r.to_string()
    ref = O.CpuPipeline(threads=1, max_inner=10)
    ref.add_cloth(V, F)
    ref.add_obstacle(SV, SF)
    contacts = first_contact = 0
    for k in range(40):
        assert w.solver_step(), k
        rc, st = ref.step()
        assert rc == 0
        s = w.last_stats
        _, x = w.get_cloth_position(h, n=V.size)
        x = x.reshape(-1, 3)
        if s.contacts == 0 and st["collisions"] == 0 and s.outer_iterations == 1:
            err = float(np.abs(x - ref.positions(0).reshape(-1, 3)).max())
            assert err <= 1e-5, (k, err)
        elif not first_contact:
            first_contact = k
        contacts += s.contacts
        assert inside_convex(x, SV, SF).min() > 0.0, k
        assert self_intersections(x, F) == 0, k
    print(f"first contact at step {first_contact}, {contacts} contacts over 40 steps")
    assert first_contact > 0 and contacts > 0


def test_cpu_backend_rejects_device_entry_points():
    w = cpu_world()
    V, F = scenes.grid_sheet(4)
    r, _ = w.add_cloth(sp.ClothConfig(), sp.CollisionConfig(), V, F, [])
    assert r
    r, _ = w.autotune()
    assert not r
