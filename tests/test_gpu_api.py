"""World API behaviour on the GPU: the C++ front end, obstacles moved through
set_obstacle_position (prescribed displacement, dcd.cu:84-95 and
ccd_fast_rejection.cu:16-28), pin targets, reset, and collision groups."""
import os

import numpy as np
import pytest

import oracle as O
import silk_physics_amd as sp
from silk_physics_amd import scenes

from helpers import coll_tuple, gpu_world, inside_convex, oracle_for, same_decisions

pytestmark = pytest.mark.gpu


def test_world_api_cpp_gpu():
    exe = os.path.join(os.path.dirname(str(sp.lib_path())), "world_api_test")
    out = os.popen(f"{exe} gpu").read()
    print(out)
    assert "FAIL" not in out
    assert out.count("PASS") >= 8


def test_rising_obstacle_lockstep_and_no_penetration():
    kw = dict(acceleration_z=-9.81, dt=1 / 60)
    V, F = scenes.grid_sheet(16, z=0.35)
    pins = scenes.top_corner_pins(16)
    SV, SF = scenes.icosphere(2, 0.3, (0.5, 0.5, 0.0))
    w = gpu_world(**kw)
    cc = sp.CollisionConfig()
    r, h = w.add_cloth(sp.ClothConfig(), cc, V, F, pins)
    r, ho = w.add_obstacle(cc, SV, SF)
    r, perm, offs = w.get_partition(h, V.shape[0])
    ow = O.OracleWorld(oracle_for(kw))
    ow.add_cloth(V, F, pins, perm, offs, coll=coll_tuple(cc))
    ow.add_obstacle(SV, SF, coll=coll_tuple(cc))
    ow.finalize()
    dt = float(np.float32(1 / 60))
    x_prev = V.astype(np.float64).reshape(-1)
    v_prev = np.zeros_like(x_prev)
    rep, free_match = [], []
    prev_pos = SV.copy()
    for i in range(30):
        lift = np.float32(0.004 * (i + 1))
        pos = SV + np.array([0, 0, lift], np.float32)
        assert w.set_obstacle_position(ho, pos)
        # The oracle steps twice from the same frame (own decisions, then the
        # GPU's): the obstacle's segment is set before each.
        ow.set_obstacle_frame(0, prev_pos, pos)
        ow.set_state(0, x_prev, v_prev)
        assert w.solver_step()
        rc, _ = ow.step()
        assert rc == 0
        _, xg = w.get_cloth_position(h, n=V.size)
        xg = xg.astype(np.float64)
        e_free = np.abs(xg - ow.positions(0)).max()
        same = same_decisions(w.decisions(), ow.decisions())
        ow.set_obstacle_frame(0, prev_pos, pos)
        ow.set_state(0, x_prev, v_prev)
        ow.set_replay(*w.decisions())
        rc, _ = ow.step()
        assert rc == 0
        assert ow.decisions()[2] == 0
        rep.append(np.abs(xg - ow.positions(0)).max())
        if same:
            free_match.append(e_free)
        s_out = inside_convex(xg, pos, SF)
        assert s_out.min() > 0.0, (i, s_out.min())
        v_prev = (xg - x_prev) / dt
        x_prev = xg
        prev_pos = pos
    print(f"rising obstacle lockstep, GPU decisions replayed: median "
          f"{np.median(rep):.3e} max {max(rep):.3e}; own decisions matched on "
          f"{len(free_match)} of 30 steps, max {max(free_match):.3e}")
    assert max(rep) <= 1e-4
    assert len(free_match) >= 15 and max(free_match) <= 1e-4


def test_pin_targets_move_pinned_vertices():
    V, F = scenes.grid_sheet(8)
    pins = scenes.top_corner_pins(8)
    w = gpu_world(acceleration_z=-9.81)
    cc = sp.CollisionConfig(is_collision_on=False, is_self_collision_on=False)
    r, h = w.add_cloth(sp.ClothConfig(), cc, V, F, pins)
    target = V[pins] + np.array([0.0, 0.0, 0.1], np.float32)
    assert w.set_cloth_pin_position(h, target)
    assert w.solver_step()
    _, x = w.get_cloth_position(h, n=V.size)
    np.testing.assert_array_equal(x.reshape(-1, 3)[pins], target)


def test_reset_restores_rest_state():
    V, F = scenes.grid_sheet(8)
    w = gpu_world()
    cc = sp.CollisionConfig(is_collision_on=False, is_self_collision_on=False)
    r, h = w.add_cloth(sp.ClothConfig(), cc, V, F, [])
    for _ in range(3):
        assert w.solver_step()
    assert w.solver_reset()
    _, x = w.get_cloth_position(h, n=V.size)
    np.testing.assert_array_equal(x, V.reshape(-1))
    assert w.solver_step()
    _, x2 = w.get_cloth_position(h, n=V.size)
    assert np.abs(x2 - V.reshape(-1)).max() > 0


@pytest.mark.parametrize("same_group", [True, False])
def test_collision_groups(same_group):
    """Two sheets, the upper one dropped on the lower, pinned one: with the
    same group the upper one stays above; with different groups it falls
    through (find_collision.cu:17-26)."""
    Vl, F = scenes.grid_sheet(10, z=0.0)
    Vu, _ = scenes.grid_sheet(10, z=0.03)
    pins_l = np.arange(Vl.shape[0], dtype=np.int32)  # lower sheet fixed
    w = gpu_world(acceleration_z=-9.81)
    r, hl = w.add_cloth(sp.ClothConfig(), sp.CollisionConfig(group=0), Vl, F, pins_l)
    r, hu = w.add_cloth(sp.ClothConfig(), sp.CollisionConfig(group=0 if same_group else 1),
                        Vu + np.float32([0.013, 0.017, 0.0]), F, [])
    for _ in range(20):
        assert w.solver_step()
    _, xu = w.get_cloth_position(hu, n=Vu.size)
    zmin = xu.reshape(-1, 3)[:, 2].min()
    if same_group:
        assert zmin > 0.0
    else:
        assert zmin < -0.05
