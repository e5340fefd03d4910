"""The headless scene driver without a device (--check): scene keys and
defaults, the four mesh formats, the transform (T Rz Ry Rx S with the
rotation taken as radians, as the reference's AffineTransformer) and the
pin box; and the trajectory reader on a constructed file."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

import silk_physics_amd as sp
from silk_physics_amd import scenes, trajectory

from scene_files import scene_json, write_obj, write_off, write_ply, write_stl

EXE = os.path.join(os.path.dirname(str(sp.lib_path())), "silk_headless")


def _check(scene):
    out = subprocess.run([EXE, "--config", str(scene), "--check"], capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    return [json.loads(l) for l in out.stdout.splitlines() if l.startswith("{")]


def _transform(V, t, r, s):
    cx, sx = np.cos(r[0]), np.sin(r[0])
    cy, sy = np.cos(r[1]), np.sin(r[1])
    cz, sz = np.cos(r[2]), np.sin(r[2])
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return (Rz @ Ry @ Rx @ (np.asarray(V, np.float64) * s).T).T + t


@pytest.mark.parametrize("fmt", ["obj", "obj_quads", "off", "ply", "ply_bin", "stl"])
def test_formats_transform_and_pins(tmp_path, fmt):
    V, F = scenes.grid_sheet(6)
    path = tmp_path / ("sheet." + fmt.split("_")[0])
    {"obj": lambda: write_obj(path, V, F), "obj_quads": lambda: write_obj(path, V, F, True),
     "off": lambda: write_off(path, V, F), "ply": lambda: write_ply(path, V, F),
     "ply_bin": lambda: write_ply(path, V, F, True),
     "stl": lambda: write_stl(path, V, F)}[fmt]()
    t, r, s = [0.5, -1.0, 2.0], [0.0, 0.0, np.pi / 2], 2.0
    W = _transform(V, t, r, s)
    # pin the row that maps to the largest x after the transform
    box = {"min": [W[:, 0].max() - 1e-3, -10, -10], "max": [10, 10, 10]}
    scene_json(tmp_path / "scene.json", [
        {"type": "cloth", "name": "sheet", "mesh": path.name,
         "cloth": {"elastic_stiffness": 150.0},
         "collision": {"enabled": True, "self_collision": True, "group": 3},
         "transform": {"translation": t, "rotation_euler_deg": r, "scale": s},
         "pin_selection": {"bbox": box}}], total_steps=7)
    objs = _check(tmp_path / "scene.json")
    o, g = objs[0], objs[1]
    assert o["name"] == "sheet" and o["cloth"] and o["group"] == 3
    assert o["vertices"] == V.shape[0] and o["triangles"] == F.shape[0]
    assert o["pins"] == 7  # one row of a 6x6-quad sheet
    np.testing.assert_allclose(o["bbox"], [*W.min(0), *W.max(0)], atol=2e-5)
    assert g["steps"] == 7 and g["max_inner_iteration"] == 10


def test_scene_defaults_and_errors(tmp_path):
    V, F = scenes.grid_sheet(2)
    write_obj(tmp_path / "m.obj", V, F)
    with open(tmp_path / "s.json", "w") as f:
        json.dump({"objects": [{"type": "obstacle", "name": "o", "mesh": "m.obj",
                                "transform": {"scale": [1.0, 2.0, 3.0]}}]}, f)
    objs = _check(tmp_path / "s.json")
    assert objs[0]["cloth"] is False and objs[0]["group"] == 0
    np.testing.assert_allclose(objs[0]["bbox"], [0, 0, 0, 1, 2, 0], atol=1e-6)
    # reference defaults (config.hpp): dt 0.01, 100 steps, 100 outer, 50 inner
    assert objs[1] == {"steps": 100, "dt": 0.01, "max_outer_iteration": 100,
                       "max_inner_iteration": 50}
    for bad in ['{"objects": [{"type": "rock", "mesh": "m.obj"}]}', '{"objects": 3}',
                '{"objects": [{"type": "cloth"}]}', '{"objects": [',
                '{"objects": [{"type": "cloth", "mesh": "missing.obj"}]}']:
        with open(tmp_path / "bad.json", "w") as f:
            f.write(bad)
        out = subprocess.run([EXE, "--config", str(tmp_path / "bad.json"), "--check"],
                             capture_output=True, text=True, timeout=60)
        assert out.returncode != 0 and out.stderr.startswith("silk_headless:"), bad


def test_trajectory_reader(tmp_path):
    rest = np.arange(9, dtype=np.float32).reshape(3, 3)
    faces = np.array([[0, 1, 2]], np.int32)
    p = tmp_path / "t.silktraj"
    with open(p, "wb") as f:
        f.write(b"SILKTRJ1" + struct.pack("<I", 2))
        for kind, moving, name in [(0, 1, b"c"), (1, 0, b"ob")]:
            f.write(struct.pack("<III", kind, moving, len(name)) + name)
            f.write(struct.pack("<II", 3, 1) + rest.tobytes() + faces.tobytes())
        f.write(struct.pack("<I", 2))
        for k in range(2):
            f.write(struct.pack("<f", 0.5 * k) + (rest + k).tobytes())
    tr = trajectory.read(str(p))
    assert tr.times.tolist() == [0.0, 0.5]
    assert [o.name for o in tr.objects] == ["c", "ob"]
    assert tr.objects[0].cloth and not tr.objects[1].moving
    np.testing.assert_array_equal(tr.objects[0].frames[1], rest + 1)
    np.testing.assert_array_equal(tr.objects[1].faces, faces)


def test_headless_cpu_backend_trajectory(tmp_path):
    """--backend cpu steps a scene file on the library's CPU solver; the same
    scene through the Python World on Backend.CPU gives every frame bit for
    bit, and no cloth vertex enters the sphere."""
    V, F = scenes.grid_sheet(10, z=0.36)
    SV, SF = scenes.icosphere(2, 0.3, (0.5, 0.5, 0.0))
    write_obj(tmp_path / "sheet.obj", V, F)
    write_obj(tmp_path / "ball.obj", SV, SF)
    coll = {"enabled": True, "self_collision": True, "group": 1}
    scene_json(tmp_path / "scene.json", [
        {"type": "cloth", "name": "sheet", "mesh": "sheet.obj", "collision": coll},
        {"type": "obstacle", "name": "ball", "mesh": "ball.obj", "collision": coll}],
        total_steps=12)
    out_file = tmp_path / "out.silktraj"
    r = subprocess.run([EXE, "--config", str(tmp_path / "scene.json"), "--backend", "cpu",
                        "--output", str(out_file), "--quiet"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    tr = trajectory.read(str(out_file))
    sheet = tr.objects[0]
    assert len(sheet.frames) == 13
    w = sp.World()
    assert w.init(sp.Backend.CPU)
    assert w.set_global_config(sp.GlobalConfig(acceleration_z=-9.81, dt=1 / 60,
                                               max_outer_iteration=100,
                                               max_inner_iteration=10))
    cc = sp.CollisionConfig(is_collision_on=True, is_self_collision_on=True, group=1)
    rr, h = w.add_cloth(sp.ClothConfig(), cc, sheet.rest, F, [])
    assert rr, rr.to_string()
    assert w.add_obstacle(cc, tr.objects[1].rest, SF)[0]
    for k in range(12):
        assert w.solver_step()
        _, x = w.get_cloth_position(h, n=V.size)
        np.testing.assert_array_equal(x.reshape(-1, 3), sheet.frames[k + 1], err_msg=str(k))
    from helpers import inside_convex
    assert inside_convex(sheet.frames[-1], SV, SF).min() > 0.0
