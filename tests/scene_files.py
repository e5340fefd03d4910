"""Writers of small mesh files and scene JSON for the headless-driver tests
(not collected by pytest)."""
from __future__ import annotations

import json
import struct

import numpy as np


def write_obj(path, V, F, quads=False):
    with open(path, "w") as f:
        for v in V:
            f.write(f"v {v[0]:.9g} {v[1]:.9g} {v[2]:.9g}\n")
        if quads:
            # pairs of triangles (a, b, c), (c, b, d) as one quad a b d c
            for k in range(0, len(F), 2):
                a, b, c = F[k]
                d = F[k + 1][2]
                f.write(f"f {a + 1}/1/1 {b + 1} {d + 1} {c + 1}\n")
        else:
            for t in F:
                f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")


def write_off(path, V, F):
    with open(path, "w") as f:
        f.write(f"OFF\n# test mesh\n{len(V)} {len(F)} 0\n")
        for v in V:
            f.write(f"{v[0]:.9g} {v[1]:.9g} {v[2]:.9g}\n")
        for t in F:
            f.write(f"3 {t[0]} {t[1]} {t[2]}\n")


def write_ply(path, V, F, binary=False):
    head = ("ply\nformat {} 1.0\nelement vertex {}\nproperty float x\nproperty float y\n"
            "property float z\nelement face {}\nproperty list uchar int vertex_indices\n"
            "end_header\n").format("binary_little_endian" if binary else "ascii", len(V), len(F))
    with open(path, "wb") as f:
        f.write(head.encode())
        if binary:
            f.write(np.asarray(V, "<f4").tobytes())
            for t in F:
                f.write(struct.pack("<Biii", 3, *map(int, t)))
        else:
            for v in V:
                f.write(f"{v[0]:.9g} {v[1]:.9g} {v[2]:.9g}\n".encode())
            for t in F:
                f.write(f"3 {t[0]} {t[1]} {t[2]}\n".encode())


def write_stl(path, V, F):
    with open(path, "wb") as f:
        f.write(b"\0" * 80)
        f.write(struct.pack("<I", len(F)))
        for t in F:
            f.write(struct.pack("<3f", 0, 0, 0))
            for k in range(3):
                f.write(np.asarray(V[t[k]], "<f4").tobytes())
            f.write(b"\0\0")


def scene_json(path, objects, **global_kw):
    g = {"dt": 1 / 60, "max_outer_iteration": 100, "max_inner_iteration": 10,
         "acceleration": [0.0, 0.0, -9.81], "total_steps": 10}
    g.update(global_kw)
    with open(path, "w") as f:
        json.dump({"global": g, "objects": objects}, f, indent=1)
