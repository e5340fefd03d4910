"""Synthetic scenes for tests and benchmarks (seeded, FP32, metres).

grid_sheet:   n x n quads over [0, span]^2 in the XY plane, split along the
              (a, b, c), (c, b, d) diagonal; optional uniform jitter.
icosphere:    subdivided icosahedron (subdiv 4 -> 2562 vertices / 5120 faces).
"""
from __future__ import annotations

import numpy as np


def grid_sheet(n: int, span: float = 1.0, z: float = 0.0, jitter: float = 0.0,
               seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    xs = np.linspace(0.0, span, n + 1, dtype=np.float64)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    V = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], axis=1)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        V = V + rng.uniform(-jitter, jitter, size=V.shape)
    a = (np.arange(n)[None, :] + (n + 1) * np.arange(n)[:, None]).ravel()
    b = a + 1
    c = a + n + 1
    d = c + 1
    F = np.stack([np.stack([a, b, c], 1), np.stack([c, b, d], 1)], 1).reshape(-1, 3)
    return V.astype(np.float32), F.astype(np.int32)


def top_corner_pins(n: int) -> np.ndarray:
    """The two vertices at y = span, x in {0, span}."""
    return np.array([n * (n + 1), n * (n + 1) + n], dtype=np.int32)


def icosphere(subdiv: int = 4, radius: float = 0.3,
              center=(0.5, 0.5, 0.0)) -> tuple[np.ndarray, np.ndarray]:
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9),
             (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2),
             (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
             (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache: dict[tuple[int, int], int] = {}

        def mid(i: int, j: int) -> int:
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius + np.asarray(center, dtype=np.float64)
    return V.astype(np.float32), np.array(faces, dtype=np.int32)


def c1_sheet() -> dict:
    """32x32 quads, 1 m, XY plane, two top corners pinned, g = -9.81 z."""
    V, F = grid_sheet(32)
    return {"V": V, "F": F, "pins": top_corner_pins(32)}


def c2_scene(n: int = 80, z: float = 0.8) -> dict:
    """n x n sheet at height z (0.5 m above the top by default) dropped onto an
    r = 0.3 icosphere (2562 vertices) at (0.5, 0.5, 0)."""
    V, F = grid_sheet(n, z=z)
    SV, SF = icosphere(4, 0.3, (0.5, 0.5, 0.0))
    return {"V": V, "F": F, "pins": np.zeros(0, np.int32), "SV": SV, "SF": SF}


def c4_scene(count: int = 64, seed: int = 0x5117B47C, nmin: int = 16, nmax: int = 128,
             sphere_subdiv: int = 3) -> list[dict]:
    """`count` independent sheets on an 8-wide lattice 3 m apart: sheet i has
    n_i ~ U{nmin..nmax} quads per side (1 m span), 0-4 random boundary
    vertices pinned, its own r = 0.3 sphere 2 cm below it and collision
    group i + 1, so no two instances can interact."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        ox, oy = 3.0 * (i % 8), 3.0 * (i // 8)
        V, F = grid_sheet(n, z=0.32)
        V[:, 0] += ox
        V[:, 1] += oy
        side = np.arange(n + 1)
        boundary = np.unique(np.concatenate([side, n * (n + 1) + side,
                                             side * (n + 1), side * (n + 1) + n]))
        pins = np.sort(rng.choice(boundary, size=int(rng.integers(0, 5)), replace=False))
        SV, SF = icosphere(sphere_subdiv, 0.3, (0.5 + ox, 0.5 + oy, 0.0))
        out.append({"V": V.astype(np.float32), "F": F, "pins": pins.astype(np.int32),
                    "SV": SV, "SF": SF, "group": i + 1, "n": n})
    return out
