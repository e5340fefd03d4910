"""The penetration checks the GPU tests use, against constructed cases."""
import numpy as np

from helpers import inside_convex, self_intersections
from silk_physics_amd import scenes


def test_self_intersections_counts_crossing_edges():
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0.5, 0.2, -0.5], [0.5, 0.8, -0.5], [0.5, 0.8, 0.5],
                  [0.5, 0.2, 0.5]], float)
    F = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7]])
    # Edges 5-6, 7-4 and the diagonal 4-6 of the vertical square pass
    # through the horizontal one (4-5 lies below it, 6-7 above), and the
    # horizontal diagonal 0-2 passes through the vertical square.
    assert self_intersections(X, F) == 4
    X[4:, 2] += 1.0
    assert self_intersections(X, F) == 0


def test_self_intersections_of_flat_and_wavy_sheets():
    V, F = scenes.grid_sheet(16, z=0.3)
    assert self_intersections(V, F) == 0
    X = V.astype(np.float64)
    X[:, 2] += 0.2 * np.sin(6 * X[:, 0])
    assert self_intersections(X, F) == 0


def test_inside_convex_signs():
    SV, SF = scenes.icosphere(2, 0.3, (0.5, 0.5, 0.0))
    s = inside_convex(np.array([[0.5, 0.5, 0.0], [0.5, 0.5, 0.31],
                                [0.5, 0.5, 0.29], [2.0, 0.5, 0.0]]), SV, SF)
    assert s[0] < -0.25 and s[1] > 0 and s[2] < 0 and s[3] > 1.0
