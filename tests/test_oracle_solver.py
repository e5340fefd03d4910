"""Oracle MAS-PCG (mas_cg_solver.cu:30-311, mas_preconditioner.cu:99-989)
against a dense solve, and the MAS preconditioner against its definition."""
import numpy as np

import oracle as O
from silk_physics_amd import scenes


def _system(n, seed):
    V, F = scenes.grid_sheet(n, jitter=0.01, seed=seed)
    V = V.astype(np.float64)
    m, _ = O.mass_area(V, F)
    h = 1 / 60
    diag = np.repeat(0.1 * m / h ** 2, 3)
    return V, F, diag


def _offsets(nv, part):
    o = list(range(0, nv, part)) + [nv]
    return np.array(sorted(set(o)), dtype=np.int32)


def test_pcg_matches_dense_solve():
    V, F, diag = _system(8, 1)
    A, _ = O.operator_dense(V, F, 100.0, 1e-4)
    Afull = A + np.diag(diag)
    rng = np.random.default_rng(0)
    b = rng.normal(size=V.size)
    x_ref = np.linalg.solve(Afull, b)
    x, info = O.pcg(V, F, 100.0, 1e-4, _offsets(V.shape[0], 30), diag, b,
                    np.zeros(V.size), 1e-8, 0.0, 2000)
    assert info["status"] == 1 and info["iters"] < V.size
    assert info["rr"] <= 1e-16 * info["rr0"]
    np.testing.assert_allclose(x, x_ref, rtol=0, atol=1e-7 * np.abs(x_ref).max())


def test_mas_single_bank_closed_form():
    # 4x4 quads = 25 vertices in one bank: level 0 is the exact inverse and
    # the single coarse node adds the Galerkin correction of the translation
    # mode, z = A^-1 r + P^T (P A P^T)^-1 P r with P summing the vertices.
    V, F, diag = _system(4, 2)
    A, _ = O.operator_dense(V, F, 100.0, 1e-4)
    Afull = A + np.diag(diag)
    r = np.random.default_rng(1).normal(size=V.size)
    z, levels = O.mas_apply(V, F, 100.0, 1e-4, np.array([0, V.shape[0]], np.int32),
                            diag, r)
    assert levels == 1
    P = np.zeros((3, V.size))
    for d in range(3):
        P[d, d::3] = 1.0
    want = np.linalg.solve(Afull, r) + P.T @ np.linalg.solve(P @ Afull @ P.T, P @ r)
    np.testing.assert_allclose(z, want, rtol=1e-9, atol=1e-12)


def test_mas_is_symmetric_positive_definite():
    V, F, diag = _system(10, 3)
    offs = _offsets(V.shape[0], 30)
    rng = np.random.default_rng(2)
    r1, r2 = rng.normal(size=V.size), rng.normal(size=V.size)
    z1, _ = O.mas_apply(V, F, 100.0, 1e-4, offs, diag, r1)
    z2, _ = O.mas_apply(V, F, 100.0, 1e-4, offs, diag, r2)
    z12, _ = O.mas_apply(V, F, 100.0, 1e-4, offs, diag, 2 * r1 - 3 * r2)
    np.testing.assert_allclose(z12, 2 * z1 - 3 * z2, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(r1 @ z2, r2 @ z1, rtol=1e-10)
    assert r1 @ z1 > 0 and r2 @ z2 > 0
