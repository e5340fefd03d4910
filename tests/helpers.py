"""Helpers shared by the test modules (not collected by pytest)."""
from __future__ import annotations

import numpy as np

import oracle as O
import silk_physics_amd as sp
from silk_physics_amd import scenes


def f32(x: float) -> float:
    """Round a Python float through FP32, as the library stores its inputs."""
    return float(np.float32(x))


def gpu_world(**global_kw) -> "sp.World":
    w = sp.World()
    r = w.init(sp.Backend.GPU)
    assert r, r.to_string()
    if global_kw:
        r = w.set_global_config(sp.GlobalConfig(**global_kw))
        assert r, r.to_string()
    return w


def oracle_for(global_kw: dict) -> dict:
    """Oracle global parameters matching GlobalConfig keyword arguments (FP32
    rounded like the library's copies)."""
    g = sp.GlobalConfig(**global_kw)
    return dict(
        ax=f32(g.acceleration_x), ay=f32(g.acceleration_y), az=f32(g.acceleration_z),
        max_outer=g.max_outer_iteration, max_inner=g.max_inner_iteration,
        dt=f32(g.dt), lin_abs=f32(g.linear_solver_abs_tol),
        lin_rel_min=f32(g.linear_solver_rel_tol_min),
        lin_rel_max=f32(g.linear_solver_rel_tol_max),
        lin_rel_init=f32(g.initial_linear_rel_tol), admm_abs=f32(g.admm_abs_tol),
        admm_rel=f32(g.admm_rel_tol), ms_scale=f32(0.2), ccd_tol=f32(1e-6),
        pcg_criterion=g.pcg_criterion, ccd_zero=g.ccd_zero_toi)


def coll_tuple(c: "sp.CollisionConfig") -> tuple:
    return (int(c.is_collision_on), int(c.is_self_collision_on), c.group,
            c.restitution, c.friction)


def sheet_on_sphere(n: int = 16, subdiv: int = 2, z: float = 0.4,
                    radius: float = 0.3):
    V, F = scenes.grid_sheet(n, z=z)
    SV, SF = scenes.icosphere(subdiv, radius, (0.5, 0.5, 0.0))
    return V, F, SV, SF


def inscribed_radius(SV: np.ndarray, SF: np.ndarray,
                     center=(0.5, 0.5, 0.0)) -> float:
    """Smallest distance from the centre to a face plane of a convex mesh."""
    c = np.asarray(center, dtype=np.float64)
    a, b, d = (SV[SF[:, k]].astype(np.float64) for k in range(3))
    n = np.cross(b - a, d - a)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return float(np.abs(np.einsum("ij,ij->i", n, a - c)).min())


def contact_keys(contacts: np.ndarray) -> set:
    """Canonical keys of a decision log's contacts (DCD call, type, vertex
    ids): a PT keeps its point and sorts the triangle, an EE sorts each edge
    and then the two edges."""
    out = set()
    for call, typ, a, b, c, d in np.asarray(contacts).reshape(-1, 6).tolist():
        if typ == 0:
            key = (a, *sorted((b, c, d)))
        else:
            e1, e2 = tuple(sorted((a, b))), tuple(sorted((c, d)))
            key = (*min(e1, e2), *max(e1, e2))
        out.add((call, typ, *key))
    return out


def same_decisions(gpu: tuple, ora: tuple, toi_rtol: float = 1e-5) -> bool:
    """GPU (TOIs, contacts, ADMM iterations, contact info) against the
    oracle's (TOIs, contacts, missing, ADMM iterations, contact info): the
    same number of outer iterations, TOIs equal to FP32 rounding, the same
    ADMM iteration counts, the same contacts found at the same DCD calls and
    the same colours and solved-vertex masks."""
    gt, gc, ga, gf = gpu[0], gpu[1], gpu[2], gpu[3]
    ot, oc, oa, of = ora[0], ora[1], ora[3], ora[4]
    if len(gt) != len(ot):
        return False
    if len(gt) and np.abs(np.asarray(gt, np.float64) - ot).max() > toi_rtol:
        return False
    if not np.array_equal(ga, oa):
        return False
    if contact_keys(gc) != contact_keys(oc):
        return False
    return np.array_equal(gf, of)


def lockstep_step(w, h, ow, nvx: int, x_prev, v_prev):
    """One GPU step and two oracle steps from the GPU's state: the oracle's
    own decisions, then the GPU's decisions replayed.  Returns (GPU
    positions, free-run error, replay error, decisions matched, replayed
    contacts the oracle could not form)."""
    ow.set_state(0, x_prev, v_prev)
    assert w.solver_step()
    _, xg = w.get_cloth_position(h, n=nvx)
    xg = xg.astype(np.float64)
    rc, _ = ow.step()
    assert rc == 0
    e_free = float(np.abs(xg - ow.positions(0)).max())
    dec_o = ow.decisions()
    dec_g = w.decisions()
    ow.set_state(0, x_prev, v_prev)
    ow.set_replay(dec_g[0], dec_g[1], dec_g[2], dec_g[3])
    rc, _ = ow.step()
    assert rc == 0
    e_rep = float(np.abs(xg - ow.positions(0)).max())
    missing = ow.decisions()[2]
    return xg, e_free, e_rep, same_decisions(dec_g, dec_o), missing


def inside_convex(x: np.ndarray, SV: np.ndarray, SF: np.ndarray) -> np.ndarray:
    """Signed distance of each point to a convex closed mesh's boundary as the
    largest face-plane distance (outward normals): negative = inside."""
    a, b, c = (SV[SF[:, k]].astype(np.float64) for k in range(3))
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    cen = SV.astype(np.float64).mean(axis=0)
    flip = np.einsum("ij,ij->i", n, a - cen) < 0
    n[flip] *= -1
    off = np.einsum("ij,ij->i", n, a)
    x = np.asarray(x, np.float64).reshape(-1, 3)
    out = np.full(x.shape[0], -np.inf)
    for s in range(0, x.shape[0], 4096):
        out[s:s + 4096] = (x[s:s + 4096] @ n.T - off).max(axis=1)
    return out


def self_intersections(x: np.ndarray, F: np.ndarray) -> int:
    """Edges of a triangle mesh that cross one of its triangles (sharing no
    vertex with the edge), by an exact segment-triangle test in FP64 over a
    uniform-grid broadphase of the triangle boxes."""
    X = np.asarray(x, np.float64).reshape(-1, 3)
    F = np.asarray(F).reshape(-1, 3)
    E = np.unique(np.sort(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]),
                          axis=1), axis=0)
    tlo = X[F].min(axis=1)
    thi = X[F].max(axis=1)
    elo = X[E].min(axis=1)
    ehi = X[E].max(axis=1)
    h = float(np.median(np.linalg.norm(X[E[:, 0]] - X[E[:, 1]], axis=1))) * 2.0
    org = X.min(axis=0)
    cell = lambda p: np.floor((p - org) / h).astype(np.int64)  # noqa: E731
    grid: dict = {}
    t0, t1 = cell(tlo), cell(thi)
    for t in range(F.shape[0]):
        for i in range(t0[t, 0], t1[t, 0] + 1):
            for j in range(t0[t, 1], t1[t, 1] + 1):
                for k in range(t0[t, 2], t1[t, 2] + 1):
                    grid.setdefault((i, j, k), []).append(t)
    e0, e1 = cell(elo), cell(ehi)
    pe, pt = [], []
    for e in range(E.shape[0]):
        cand = set()
        for i in range(e0[e, 0], e1[e, 0] + 1):
            for j in range(e0[e, 1], e1[e, 1] + 1):
                for k in range(e0[e, 2], e1[e, 2] + 1):
                    cand.update(grid.get((i, j, k), ()))
        for t in cand:
            pe.append(e)
            pt.append(t)
    if not pe:
        return 0
    pe, pt = np.array(pe), np.array(pt)
    share = (F[pt] == E[pe, 0:1]).any(axis=1) | (F[pt] == E[pe, 1:2]).any(axis=1)
    box = ((elo[pe] <= thi[pt]) & (ehi[pe] >= tlo[pt])).all(axis=1)
    keep = ~share & box
    pe, pt = pe[keep], pt[keep]
    p, q = X[E[pe, 0]], X[E[pe, 1]]
    a, b, c = X[F[pt, 0]], X[F[pt, 1]], X[F[pt, 2]]
    # Moller-Trumbore on the segment p + s (q - p), s in [0, 1].
    d = q - p
    e1v, e2v = b - a, c - a
    pv = np.cross(d, e2v)
    det = np.einsum("ij,ij->i", e1v, pv)
    ok = np.abs(det) > 1e-30
    inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
    tv = p - a
    u = np.einsum("ij,ij->i", tv, pv) * inv
    qv = np.cross(tv, e1v)
    v = np.einsum("ij,ij->i", d, qv) * inv
    s = np.einsum("ij,ij->i", e2v, qv) * inv
    hit = ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (s >= 0) & (s <= 1)
    return int(hit.sum())
