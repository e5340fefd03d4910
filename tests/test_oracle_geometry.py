"""Oracle geometry and operator against closed forms and invariants
(mass: Voronoi areas of libigl massmatrix VORONOI; Jacobian:
cloth_assembly_l2_cache.cpp:27-69; operator: cloth_assembly_l2_cache.cpp:
71-137 with k JWJ + kb CWC)."""
import numpy as np

import oracle as O
from silk_physics_amd import scenes


def test_right_triangle_voronoi_mass():
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float64)
    F = np.array([[0, 1, 2]])
    m, a = O.mass_area(V, F)
    assert a[0] == 0.5
    # right angle at vertex 0: circumcentre on the hypotenuse midpoint
    np.testing.assert_allclose(m, [0.25, 0.125, 0.125], rtol=1e-14)


def test_obtuse_triangle_mass_rule():
    V = np.array([[0, 0, 0], [2, 0, 0], [1, 0.2, 0]], dtype=np.float64)
    m, a = O.mass_area(V, np.array([[0, 1, 2]]))
    # obtuse at vertex 2: 1/2 of the area there, 1/4 at the others
    np.testing.assert_allclose(m, [a[0] / 4, a[0] / 4, a[0] / 2], rtol=1e-14)


def test_mass_sums_to_area():
    V, F = scenes.grid_sheet(8, jitter=0.01, seed=3)
    m, a = O.mass_area(V.astype(np.float64), F)
    np.testing.assert_allclose(m.sum(), a.sum(), rtol=1e-13)
    np.testing.assert_allclose(a.sum(), 1.0, atol=2e-2)


def _deformation(J, X, F):
    out = []
    for f in range(F.shape[0]):
        x = X[F[f]].reshape(-1)
        y = J[f] @ x
        out.append(np.stack([y[:3], y[3:]], axis=1))  # 3x2, columns
    return np.array(out)


def test_jacobian_rest_is_orthonormal_and_scales():
    V, F = scenes.grid_sheet(4, jitter=0.02, seed=1)
    V = V.astype(np.float64)
    J = O.jacobians(V, F)
    D = _deformation(J, V, F)
    eye = np.eye(2)
    for Fi in D:
        np.testing.assert_allclose(Fi.T @ Fi, eye, atol=1e-12)
    D2 = _deformation(J, 1.7 * V, F)
    for Fi in D2:
        np.testing.assert_allclose(Fi.T @ Fi, 1.7 ** 2 * eye, atol=1e-12)
    # rigid rotation keeps F^T F = I
    th = 0.7
    R = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    D3 = _deformation(J, V @ R.T + 0.3, F)
    for Fi in D3:
        np.testing.assert_allclose(Fi.T @ Fi, eye, atol=1e-12)


def test_jacobian_annihilates_translation():
    V, F = scenes.grid_sheet(3)
    J = O.jacobians(V.astype(np.float64), F)
    t = np.tile([0.3, -1.0, 2.0], 3)
    assert np.abs(J @ t).max() < 1e-12


def test_operator_symmetric_psd_translation_free():
    V, F = scenes.grid_sheet(4, jitter=0.01, seed=2)
    A, b = O.operator_dense(V.astype(np.float64), F, 100.0, 1e-4)
    np.testing.assert_allclose(A, A.T, atol=1e-10)
    ev = np.linalg.eigvalsh(A)
    assert ev.min() > -1e-9 * ev.max()
    # translations are in the null space (J and C annihilate constants)
    for d in range(3):
        t = np.zeros(A.shape[0])
        t[d::3] = 1.0
        assert np.abs(A @ t).max() < 1e-9 * np.abs(A).max()


def test_operator_quadratic_form_is_stretch_energy():
    # k JWJ part: x^T A x = k sum_f area_f ||F_f(x)||_F^2 (kb = 0)
    V, F = scenes.grid_sheet(3, jitter=0.01, seed=5)
    V = V.astype(np.float64)
    k = 100.0
    A, _ = O.operator_dense(V, F, k, 0.0)
    rng = np.random.default_rng(0)
    x = rng.normal(size=V.size)
    J = O.jacobians(V, F)
    _, area = O.mass_area(V, F)
    e = sum(area[f] * np.sum((J[f] @ x.reshape(-1, 3)[F[f]].reshape(-1)) ** 2)
            for f in range(F.shape[0]))
    np.testing.assert_allclose(x @ A @ x, k * e, rtol=1e-12)


def test_project_stretch_closed_forms():
    # identity-like F (orthonormal columns) is already feasible
    y = np.array([1.0, 0, 0, 0, 1.0, 0])
    np.testing.assert_allclose(O.project_stretch(y), y, atol=1e-14)
    # uniform stretch s: sigma clamps to [0.9, 1.1]
    for s, c in [(1.5, 1.1), (0.5, 0.9), (1.05, 1.05)]:
        np.testing.assert_allclose(O.project_stretch(s * y), c * y, atol=1e-14)
    # rotation invariance
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    Fm = rng.normal(size=(3, 2))
    p = O.project_stretch(Fm.T.reshape(-1)).reshape(2, 3).T
    pr = O.project_stretch((Q @ Fm).T.reshape(-1)).reshape(2, 3).T
    np.testing.assert_allclose(Q @ p, pr, atol=1e-12)


def test_project_stretch_is_nearest_feasible():
    # brute force: the projection minimises ||F - U diag(s') V^T|| over
    # s' in [0.9, 1.1]^2 with the SVD frames of F
    rng = np.random.default_rng(7)
    for _ in range(20):
        Fm = rng.normal(size=(3, 2))
        U, S, Vt = np.linalg.svd(Fm, full_matrices=False)
        p = O.project_stretch(Fm.T.reshape(-1)).reshape(2, 3).T
        want = U @ np.diag(np.clip(S, 0.9, 1.1)) @ Vt
        np.testing.assert_allclose(p, want, atol=1e-10)
        grid = np.linspace(0.9, 1.1, 21)
        best = min(np.linalg.norm(Fm - U @ np.diag([a, b]) @ Vt)
                   for a in grid for b in grid)
        assert np.linalg.norm(Fm - p) <= best + 1e-12
