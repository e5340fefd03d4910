"""Batched instances: sheets of uneven size (16..128 quads per side, seeded),
each over its own sphere with 0-4 random boundary pins and its own collision
group, stepped together in one World (one PCG system per sheet in every
solve launch).  Every solve meets the true-residual test, no sheet enters
its sphere or itself, pinned vertices stay at their rest positions, and no
sheet interacts with another (each stays over its own lattice cell).  The
ADMM loop runs with the reference's defaults (up to 100 iterations, stopped
by its convergence test); cut to 10 iterations, sheets pressed against their
spheres cross them from the 6th step on (see the 128x128 test)."""
import numpy as np
import pytest

import silk_physics_amd as sp
from silk_physics_amd import scenes

from helpers import gpu_world, inside_convex, self_intersections

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("count,steps", [(8, 12)])
def test_c4_batched_sheets(count, steps):
    sheets = scenes.c4_scene(count)
    w = gpu_world(acceleration_z=-9.81, dt=1 / 60,
                  linear_solver_rel_tol_min=1e-6, linear_solver_rel_tol_max=1e-6,
                  initial_linear_rel_tol=1e-6, pcg_criterion=sp.PCG_TRUE)
    hs = []
    for s in sheets:
        cc = sp.CollisionConfig(group=s["group"])
        r, h = w.add_cloth(sp.ClothConfig(), cc, s["V"], s["F"], s["pins"])
        assert r, r.to_string()
        r, _ = w.add_obstacle(cc, s["SV"], s["SF"])
        assert r, r.to_string()
        hs.append(h)
    info = w.scene_info()
    print(f"C4 x{count}: {info}")
    assert info["vertices"] == sum(s["V"].shape[0] for s in sheets)
    its = sols = 0
    for k in range(steps):
        assert w.solver_step(), k
        st = w.last_stats
        assert st.pcg_unverified_solves == 0 and st.pcg_stalled_solves == 0
        its += st.pcg_iterations
        sols += st.pcg_solves
    worst_in = np.inf
    for s, h in zip(sheets, hs):
        _, x = w.get_cloth_position(h, n=s["V"].size)
        x = x.reshape(-1, 3)
        worst_in = min(worst_in, inside_convex(x, s["SV"], s["SF"]).min())
        assert inside_convex(x, s["SV"], s["SF"]).min() > 0.0
        assert self_intersections(x, s["F"]) == 0
        np.testing.assert_array_equal(x[s["pins"]], s["V"][s["pins"]])
        lo, hi = s["V"].min(0), s["V"].max(0)
        assert (x[:, :2] > lo[:2] - 0.5).all() and (x[:, :2] < hi[:2] + 0.5).all()
    print(f"{steps} steps: {its / max(1, sols):.1f} PCG iterations per solve, closest "
          f"vertex {worst_in:.3e} m outside its sphere")


def test_persistent_batches_match_one_launch(monkeypatch):
    """The same 8 sheets solved with every system in one persistent launch
    and with the systems split over several launches (at most 64
    workgroups each): systems never exchange data, so the positions agree
    bit for bit."""
    def run():
        sheets = scenes.c4_scene(8)
        w = gpu_world(acceleration_z=-9.81, dt=1 / 60, max_inner_iteration=10,
                      linear_solver_rel_tol_min=1e-6, linear_solver_rel_tol_max=1e-6,
                      initial_linear_rel_tol=1e-6, pcg_criterion=sp.PCG_TRUE)
        hs = []
        for s in sheets:
            cc = sp.CollisionConfig(group=s["group"])
            r, h = w.add_cloth(sp.ClothConfig(), cc, s["V"], s["F"], s["pins"])
            assert r
            r, _ = w.add_obstacle(cc, s["SV"], s["SF"])
            assert r
            hs.append((h, s["V"].size))
        info = w.scene_info()
        for _ in range(6):
            assert w.solver_step()
        return info, [w.get_cloth_position(h, n=n)[1] for h, n in hs]

    monkeypatch.delenv("SILK_HIP_PCG_BATCH_WG", raising=False)
    info1, x1 = run()
    monkeypatch.setenv("SILK_HIP_PCG_BATCH_WG", "64")
    info2, x2 = run()
    print(f"one batch: {info1['pcg_launches']} launch(es) of {info1['pcg_workgroups']} "
          f"workgroups; capped: {info2['pcg_launches']} launches")
    assert info1["pcg_workgroups"] > 0 and info2["pcg_launches"] > info1["pcg_launches"]
    for a, b in zip(x1, x2):
        np.testing.assert_array_equal(a, b)
