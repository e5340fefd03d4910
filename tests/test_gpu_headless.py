"""The headless driver on the GPU: a pinned 12x12 sheet over a rotating
sphere from a scene file, 20 steps with the trajectory written; the same
scene replayed through the Python World API from the trajectory's own load
positions and obstacle frames reproduces every cloth frame, and the --bench
mode runs without output."""
import os
import subprocess

import numpy as np
import pytest

import silk_physics_amd as sp
from silk_physics_amd import scenes, trajectory

from helpers import gpu_world
from scene_files import scene_json, write_obj

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(str(sp.lib_path())), "silk_headless")


def test_headless_trajectory_matches_world_api(tmp_path):
    V, F = scenes.grid_sheet(12, z=0.36)
    SV, SF = scenes.icosphere(2, 0.3, (0.5, 0.5, 0.0))
    write_obj(tmp_path / "sheet.obj", V, F)
    write_obj(tmp_path / "ball.obj", SV, SF)
    coll = {"enabled": True, "self_collision": True, "group": 1, "restitution": 0.0,
            "friction": 0.4}
    scene_json(tmp_path / "scene.json", [
        {"type": "cloth", "name": "sheet", "mesh": "sheet.obj",
         "cloth": {"elastic_stiffness": 120.0, "bending_stiffness": 2e-4,
                   "density": 0.1, "damping": 0.01}, "collision": coll,
         "pin_selection": {"bbox": {"min": [-0.01, 0.99, 0.0], "max": [0.01, 1.01, 1.0]}}},
        {"type": "obstacle", "name": "ball", "mesh": "ball.obj", "collision": coll,
         "angular_velocity_deg_s": [0.0, 0.0, 90.0]}], total_steps=20)
    out_file = tmp_path / "out.silktraj"
    r = subprocess.run([EXE, "--config", str(tmp_path / "scene.json"), "--output",
                        str(out_file), "--quiet"], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    tr = trajectory.read(str(out_file))
    assert len(tr.times) == 21
    sheet, ball = tr.objects
    assert sheet.cloth and ball.moving and len(ball.frames) == 21
    # The sphere turned by 20 * 1.5 degrees about its vertex mean.
    c = ball.rest.mean(axis=0)
    a0 = np.arctan2(ball.rest[:, 1] - c[1], ball.rest[:, 0] - c[0])
    a1 = np.arctan2(ball.frames[-1][:, 1] - c[1], ball.frames[-1][:, 0] - c[0])
    moved = np.linalg.norm(ball.rest[:, :2] - c[:2], axis=1) > 0.05
    turn = np.rad2deg(np.angle(np.exp(1j * (a1 - a0))))[moved]
    np.testing.assert_allclose(turn, 30.0, atol=1e-3)

    w = gpu_world(acceleration_z=-9.81, dt=1 / 60, max_outer_iteration=100,
                  max_inner_iteration=10)
    cc = sp.CollisionConfig(is_collision_on=True, is_self_collision_on=True, group=1,
                            restitution=0.0, friction=0.4)
    pins = np.flatnonzero((np.abs(sheet.rest[:, 0]) <= 0.01) &
                          (np.abs(sheet.rest[:, 1] - 1.0) <= 0.01)).astype(np.int32)
    assert pins.size == 1
    r, h = w.add_cloth(sp.ClothConfig(elastic_stiffness=120.0, bending_stiffness=2e-4,
                                      density=0.1, damping=0.01), cc, sheet.rest,
                       sheet.faces, pins)
    assert r
    r, hb = w.add_obstacle(cc, ball.rest, ball.faces)
    assert r
    worst = 0.0
    for k in range(20):
        assert w.set_obstacle_position(hb, ball.frames[k + 1])
        assert w.solver_step()
        _, x = w.get_cloth_position(h, n=sheet.rest.size)
        worst = max(worst, float(np.abs(x.reshape(-1, 3) - sheet.frames[k + 1]).max()))
    print(f"headless vs World API over 20 steps: max |dx| {worst:.3e} m")
    assert worst <= 1e-6

    r = subprocess.run([EXE, "--config", str(tmp_path / "scene.json"), "--bench",
                        "--steps", "5", "--quiet"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    assert "5 steps" in r.stdout
