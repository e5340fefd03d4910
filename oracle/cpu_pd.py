"""The reference CPU solver step for scenes without contacts, restated in
NumPy/SciPy in FP64 (TEST INFRASTRUCTURE ONLY: tests/ hold the library's
CPU backend to it; nothing in the product imports it).

Follows silk/src/backend/cpu/solver/pipeline.cpp:45-168 with no active
collision (the CCD finds no impact, so one outer iteration takes the whole
step): velocities damped by (1 - damping) (pipeline.cpp:205-219), the
prediction x + h v + h^2 g, the outer right-hand side
M/h^2 x + M/h v + M g + kb C0 + penalty * pin targets
(cloth_solver_utils.cpp:101-163), inner iterations of per-face projections
(singular values of the 3x2 deformation clamped to [0.9, 1.1]) and global
solves with H = M/h^2 + k JWJ + kb CWC + penalty on pinned vertices
(cloth_solver_context.cpp:21-73, cloth_solver_utils.cpp:205-294) until the
update is at most 5 % of the scene's bounding-box diagonal
(pipeline.cpp:96-121), then v = (x_new - x) / h.

Geometry (Voronoi masses, areas, per-face Jacobians, k JWJ + kb CWC, kb C0)
comes from the oracle's FP64 C++ functions, which tests pin by closed forms;
the projection is NumPy's batched SVD, pinned against the oracle's
project_stretch.  The sparse system is factored once by SciPy's SuperLU.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sps
import scipy.sparse.linalg as spla

from . import jacobians, mass_area, operator_dense

PIN_PENALTY = 1e4  # cloth_solver_context.cpp:17


def project_batch(D: np.ndarray, lo: float = 0.9, hi: float = 1.1) -> np.ndarray:
    """Rows of D are 3x2 deformations stored column by column (J x of one
    face); returns U clamp(S) V^T in the same layout."""
    Fm = D.reshape(-1, 2, 3).transpose(0, 2, 1)
    U, s, Vt = np.linalg.svd(Fm, full_matrices=False)
    T = U @ (np.clip(s, lo, hi)[..., None] * Vt)
    return T.transpose(0, 2, 1).reshape(-1, 6)


class CpuPD:
    """Steps one cloth with the reference CPU algorithm (no collision)."""

    def __init__(self, V, F, pins=(), dt=1 / 60, g=(0.0, 0.0, -9.81), k=100.0,
                 kb=1e-4, density=0.1, damping=0.01, max_inner=100):
        V = np.asarray(V, dtype=np.float64).reshape(-1, 3)
        self.F = np.asarray(F, dtype=np.int32).reshape(-1, 3)
        self.nv = V.shape[0]
        self.h = float(dt)
        self.g = np.asarray(g, dtype=np.float64)
        self.k = float(k)
        self.damping = float(damping)
        self.max_inner = int(max_inner)
        mass, self.area = mass_area(V, self.F)
        self.M = density * mass
        self.J = jacobians(V, self.F)  # (nf, 6, 9)
        A, c0 = operator_dense(V, self.F, k, kb)
        self.pins = np.asarray(pins, dtype=np.int64).reshape(-1)
        diag = np.repeat(self.M / self.h ** 2, 3)
        for p in self.pins:
            diag[3 * p:3 * p + 3] += PIN_PENALTY
        H = sps.csc_matrix(A) + sps.diags(diag)
        self.lu = spla.splu(H.tocsc())
        self.rest_rhs = c0.copy()
        self.pin_pos = V[self.pins].copy()
        self.x = V.reshape(-1).copy()
        self.v = np.zeros_like(self.x)
        self.last_inner = 0

    def _elastic_rhs(self, x: np.ndarray) -> np.ndarray:
        X = x.reshape(-1, 3)[self.F].reshape(-1, 9)          # (nf, 9)
        D = np.einsum("frc,fc->fr", self.J, X)               # (nf, 6)
        T = project_batch(D)
        w = self.k * self.area
        loc = np.einsum("frc,fr->fc", self.J, T) * w[:, None]  # (nf, 9)
        out = np.zeros(3 * self.nv)
        np.add.at(out.reshape(-1, 3), self.F.reshape(-1), loc.reshape(-1, 3))
        return out

    def step(self) -> None:
        h = self.h
        x = self.x
        vel = (1.0 - self.damping) * self.v
        X = x.reshape(-1, 3)
        scale = float(np.linalg.norm(X.max(0) - X.min(0)))
        m3 = np.repeat(self.M, 3)
        acc = np.tile(self.g, self.nv)
        rhs0 = self.rest_rhs.copy()
        for i, p in enumerate(self.pins):
            rhs0[3 * p:3 * p + 3] += PIN_PENALTY * self.pin_pos[i]
        outer = rhs0 + m3 / h ** 2 * x + m3 / h * vel + m3 * acc
        nxt = x + h * vel + h * h * acc
        self.last_inner = 0
        for _ in range(self.max_inner):
            self.last_inner += 1
            sol = self.lu.solve(outer + self._elastic_rhs(nxt))
            done = np.linalg.norm(sol - nxt) <= 0.05 * scale
            nxt = sol
            if done:
                break
        self.v = (nxt - x) / h
        self.x = nxt
