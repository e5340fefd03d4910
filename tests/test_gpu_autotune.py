"""Launch autotuning (silk_hip_autotune): it needs one step on the scene,
reports a time for every broadphase split depth and PCG variant it tried,
and leaves the trajectory bit for bit as it was (the candidate lists are
sorted before use and both PCG variants run the same arithmetic; the entry
point itself checks that the two variants' solutions agree)."""
import numpy as np
import pytest

import silk_physics_amd as sp
from silk_physics_amd import scenes

from helpers import gpu_world

pytestmark = pytest.mark.gpu

KW = dict(acceleration_z=-9.81, dt=1 / 60, max_inner_iteration=10,
          linear_solver_rel_tol_min=1e-6, linear_solver_rel_tol_max=1e-6,
          initial_linear_rel_tol=1e-6)


def _world(s):
    w = gpu_world(**KW)
    cc = sp.CollisionConfig()
    r, h = w.add_cloth(sp.ClothConfig(), cc, s["V"], s["F"], s["pins"])
    assert r, r.to_string()
    r, _ = w.add_obstacle(cc, s["SV"], s["SF"])
    assert r, r.to_string()
    return w, h


def test_autotune_needs_a_step():
    s = scenes.c2_scene(24, z=0.32)
    w, _ = _world(s)
    r, rep = w.autotune()
    assert not r
    assert rep["split_depth"] == -1


def test_autotuned_steps_are_bit_identical():
    s = scenes.c2_scene(48, z=0.31)
    plain, hp = _world(s)
    tuned, ht = _world(s)
    n = s["V"].size
    for w in (plain, tuned):
        r = w.solver_step()
        assert r, r.to_string()
    r, rep = tuned.autotune()
    print(f"autotune report: {rep}")
    assert r, r.to_string()
    assert 2 <= rep["split_depth"] <= 7
    assert rep["broadphase_us"][rep["split_depth"]] == min(rep["broadphase_us"].values())
    assert all(t >= 0 for t in rep["broadphase_us"].values())
    assert rep["pcg_variant"] in (0, 1)
    assert rep["pcg_us"][rep["pcg_variant"]] > 0
    contacts = 0
    for i in range(8):
        for w in (plain, tuned):
            r = w.solver_step()
            assert r, f"step {i}: {r.to_string()}"
        contacts += tuned.last_stats.contacts
        _, xp = plain.get_cloth_position(hp, n=n)
        _, xt = tuned.get_cloth_position(ht, n=n)
        assert np.array_equal(xp, xt), f"step {i}"
    assert contacts > 0
