"""Closed forms and invariants that pin the oracle's world-level functions:
DCD records and Daviet contact assembly (dcd.cu:102-394,
contact_constraints.cu:23-150), the contact weight's smallest eigenvalue
(cloth_assembly_l1_cache.cu:60-78), contact colouring
(contact_coloring.cu:69-373), the contact Gauss-Seidel sweeps with the
Signorini-Coulomb conditions they must satisfy (contact_constraints.cu:
203-411), the ADMM local step (cloth_admm_helper.cu:81-201), the broadphase
candidate filters (find_collision.cu:17-138) and the product's vertex
partition (graph_partition_test.cpp:11-30)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import silk_physics_amd as sp
from silk_physics_amd import scenes


def _rot(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


# ------------------------------------------------------------------ DCD


@pytest.mark.parametrize("seed", range(5))
def test_pt_point_above_centroid(seed):
    """A point at height h above a triangle's centroid: distance h, u = v =
    1/3, the normal is the triangle's (towards the point), and the contact's
    affine map gives u(x) = (h - dhat) n at the detection positions."""
    rng = np.random.default_rng(seed)
    R = _rot(rng)
    tri = np.array([[0, 0, 0], [1.0, 0.1, 0], [0.3, 0.8, 0]]) * rng.uniform(0.01, 1)
    h, dhat = 0.2 * rng.uniform(0.001, 0.01), 0.01
    c = tri.mean(axis=0)
    x = np.vstack([c + [0, 0, h], tri]) @ R.T + rng.normal(size=3)
    inv = np.array([1.0, 2.0, 0.5, 1.5])
    cw = np.array([3.0, 1.0, 2.0, 4.0])
    r = O.dcd_contact(True, x, inv, [0, 0, 0, 0], cw, dhat)
    assert r is not None
    n = R @ np.array([0, 0, 1.0])
    np.testing.assert_allclose(r["dist"], h, rtol=1e-9)
    np.testing.assert_allclose([r["u"], r["v"]], [1 / 3, 1 / 3], atol=1e-9)
    np.testing.assert_allclose(r["normal"], n, atol=1e-9)
    np.testing.assert_allclose(r["coef"], [1, -1 / 3, -1 / 3, -1 / 3], atol=1e-9)
    np.testing.assert_allclose(r["u_at_x"], (h - dhat) * n, atol=1e-12)
    # The correction scatter moves u by exactly the correction:
    # sum_k coef_k gamma_k = 1, and gamma_k cw_k is proportional to coef_k.
    np.testing.assert_allclose(np.sum(r["coef"] * r["gamma"]), 1.0, rtol=1e-12)
    ratio = r["gamma"] * cw / r["coef"]
    np.testing.assert_allclose(ratio, ratio[0], rtol=1e-12)
    # Out of reach or outside the triangle: no contact.
    far = x.copy()
    far[0] = (c + [0, 0, 2 * dhat]) @ R.T + (x[1] - tri[0] @ R.T)
    assert O.dcd_contact(True, far, inv, [0, 0, 0, 0], cw, dhat) is None
    out = x.copy()
    out[0] = (tri[0] - 0.5 * (c - tri[0]) + [0, 0, h]) @ R.T + (x[1] - tri[0] @ R.T)
    assert O.dcd_contact(True, out, inv, [0, 0, 0, 0], cw, dhat) is None


def test_pt_static_and_pinned_vertices():
    """Static (obstacle) vertices carry no state and their terms sit in the
    affine part; pinned vertices keep a state but get no correction; u(x) is
    unchanged by either."""
    tri = np.array([[0, 0, 0], [1.0, 0, 0], [0, 1.0, 0]])
    h, dhat = 0.004, 0.01
    x = np.vstack([[0.25, 0.25, h], tri])
    inv = np.array([1.0, 1.0, 0.0, 1.0])  # tri1 pinned
    cw = np.ones(4)
    r = O.dcd_contact(True, x, inv, [0, 1, 0, 0], cw, dhat)  # tri0 static
    assert r is not None
    assert r["state"].tolist() == [True, False, True, True]
    assert r["gamma"][1] == 0.0 and r["gamma"][2] == 0.0
    np.testing.assert_allclose(np.sum(r["coef"] * r["gamma"]), 1.0, rtol=1e-12)
    np.testing.assert_allclose(r["u_at_x"], [0, 0, h - dhat], atol=1e-15)
    # Nothing can move: no contact.
    assert O.dcd_contact(True, x, [0, 0, 0, 0], [0, 0, 0, 0], cw, dhat) is None


def test_ee_crossing_and_parallel_edges():
    dhat = 0.01
    h = 0.003
    # Perpendicular crossing edges: midpoints, normal from edge b to edge a.
    x = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, h], [0, 1, h]], dtype=float)
    r = O.dcd_contact(False, x, np.ones(4), [0] * 4, np.ones(4), dhat)
    np.testing.assert_allclose([r["dist"], r["u"], r["v"]], [h, 0.5, 0.5], atol=1e-12)
    np.testing.assert_allclose(r["normal"], [0, 0, -1], atol=1e-12)
    np.testing.assert_allclose(r["u_at_x"], [0, 0, -(h - dhat)], atol=1e-12)
    # Parallel overlapping edges (dcd.cu:242-315 fallback): the distance is
    # still h and the pair found is an endpoint and its projection.
    x = np.array([[0, 0, 0], [1, 0, 0], [0.5, 0, h], [1.5, 0, h]], dtype=float)
    r = O.dcd_contact(False, x, np.ones(4), [0] * 4, np.ones(4), dhat)
    np.testing.assert_allclose(r["dist"], h, rtol=1e-12)
    u, v = r["u"], r["v"]
    assert min(abs(u), abs(u - 1), abs(v), abs(v - 1)) < 1e-12
    pa = (1 - u) * x[0] + u * x[1]
    pb = (1 - v) * x[2] + v * x[3]
    np.testing.assert_allclose(np.dot(pb - pa, x[1] - x[0]), 0.0, atol=1e-12)


# ------------------------------------------------------------ weights


@pytest.mark.parametrize("seed", range(4))
def test_sym3_min_eig_matches_analytic(seed):
    rng = np.random.default_rng(seed)
    for _ in range(50):
        Q = _rot(rng)
        lam = rng.uniform(-2, 5, 3)
        if rng.random() < 0.3:
            lam[1] = lam[0]  # repeated eigenvalue
        A = Q @ np.diag(lam) @ Q.T
        assert abs(O.sym3_min_eig(A) - lam.min()) <= 1e-11 * max(1, np.abs(lam).max())
    # kron(s, I3) + m I3, the isotropic block of the PD operator.
    assert O.sym3_min_eig(np.eye(3) * 7.5) == pytest.approx(7.5, rel=1e-15)


# ------------------------------------------------------------ contacts


def _random_contacts(rng, nc, nv, share=True):
    state = np.full((nc, 4), -1, dtype=np.int32)
    coef = np.zeros((nc, 4))
    gamma = np.zeros((nc, 4))
    pool = np.arange(nv)
    used = 0
    for c in range(nc):
        k = rng.integers(2, 5)
        if share:
            vs = rng.choice(pool, size=k, replace=False)
        else:
            vs = pool[used:used + k]
            used += k
        state[c, :k] = vs
        coef[c, :k] = rng.uniform(-1, 1, k)
        movable = rng.random(k) < 0.85
        movable[0] = True
        gamma[c, :k] = np.where(movable, coef[c, :k], 0.0)
        gamma[c] /= np.sum(coef[c] * gamma[c])
    normal = rng.normal(size=(nc, 3))
    normal /= np.linalg.norm(normal, axis=1, keepdims=True)
    affine = rng.normal(size=(nc, 3)) * 0.01
    mu = rng.uniform(0, 0.8, nc)
    return state, coef, gamma, normal, affine, mu


@pytest.mark.parametrize("seed", range(3))
def test_colouring_is_first_fit_over_movable_vertices(seed):
    rng = np.random.default_rng(seed)
    nv, nc = 40, 120
    st, coef, gamma, nrm, aff, mu = _random_contacts(rng, nc, nv)
    x = rng.normal(size=3 * nv)
    _, _, _, col = O.contact_sweeps(st, coef, gamma, nrm, aff, mu, x, np.ones(nv),
                                    sweeps=0)
    assert col.min() >= 1
    movable = [set(st[c][gamma[c] != 0].tolist()) for c in range(nc)]
    for c in range(nc):
        # No two contacts of one colour share a movable vertex ...
        for o in range(c):
            if col[o] == col[c]:
                assert not (movable[o] & movable[c]), (o, c)
        # ... and each takes the smallest colour its earlier neighbours leave.
        taken = {col[o] for o in range(c) if movable[o] & movable[c]}
        first = 1
        while first in taken:
            first += 1
        assert col[c] == first


def _u_of(st, coef, aff, p):
    u = aff.copy()
    for c in range(st.shape[0]):
        for k in range(4):
            if st[c, k] >= 0:
                u[c] += coef[c, k] * p[3 * st[c, k]:3 * st[c, k] + 3]
    return u


def _check_signorini_coulomb(u, force, n, mu, tol):
    un = u @ n
    fn = force @ n
    ft = force - fn * n
    assert un >= -tol                                   # no penetration
    assert fn >= -tol                                   # pushes only
    assert np.linalg.norm(ft) <= mu * fn + tol          # inside the cone
    assert abs(un * fn) <= tol * max(1.0, abs(fn))      # complementarity
    ut = u - un * n
    if np.linalg.norm(ft) > tol and np.linalg.norm(ut) > tol:
        # sliding: friction opposes the tangential motion
        assert np.dot(ft, ut) <= tol * np.linalg.norm(ut)


@pytest.mark.parametrize("seed", range(3))
def test_isolated_contacts_satisfy_signorini_coulomb_after_one_sweep(seed):
    """Contacts with disjoint vertices: one sweep is the exact projection of
    each contact, so the new u and the force (its change) satisfy the
    Signorini-Coulomb conditions exactly."""
    rng = np.random.default_rng(seed)
    nc = 30
    nv = 4 * nc
    st, coef, gamma, nrm, aff, mu = _random_contacts(rng, nc, nv, share=False)
    x = rng.normal(size=3 * nv) * 0.01
    cw = rng.uniform(0.5, 2, nv)
    u0 = _u_of(st, coef, aff, x)
    p, eta, corr, col = O.contact_sweeps(st, coef, gamma, nrm, aff, mu, x, cw, sweeps=1)
    assert col.max() == 1
    u1 = _u_of(st, coef, aff, p)
    for c in range(nc):
        np.testing.assert_allclose(u1[c] - u0[c], corr[c], atol=1e-12)
        _check_signorini_coulomb(u1[c], u1[c] - u0[c], nrm[c], mu[c], 1e-12)
    # The multiplier update eta += w (p - x) on the active vertices.
    act = np.zeros(nv, bool)
    for c in range(nc):
        act[st[c][(st[c] >= 0) & (gamma[c] != 0)]] = True
    w3 = np.repeat(np.where(act, cw, 0.0), 3)
    np.testing.assert_allclose(eta, w3 * (p - x), atol=1e-15)


def test_coupled_contacts_converge_to_signorini_coulomb():
    """Contacts sharing vertices: Gauss-Seidel sweeps converge to a point
    where every contact satisfies the conditions with its accumulated
    correction as the force."""
    rng = np.random.default_rng(11)
    nv, nc = 30, 25
    st, coef, gamma, nrm, aff, mu = _random_contacts(rng, nc, nv)
    x = rng.normal(size=3 * nv) * 0.01
    cw = np.ones(nv)
    p, eta, corr, col = O.contact_sweeps(st, coef, gamma, nrm, aff, mu, x, cw,
                                         sweeps=3000)
    assert col.max() > 1
    u = _u_of(st, coef, aff, p)
    for c in range(nc):
        _check_signorini_coulomb(u[c], corr[c], nrm[c], mu[c], 1e-7)


# ---------------------------------------------------------- local step


@pytest.mark.parametrize("stretch", [1.0, 1.2, 0.8])
def test_local_step_uniform_stretch(stretch):
    """Rest shape scaled by s: F = s R, the projection is clamp(s) R, z is the
    midpoint (rho = k), u = w rho (s - z_s) R, so the sums are closed forms in
    the total area."""
    V, F = scenes.grid_sheet(6, jitter=0.0)
    V = V.astype(np.float64)
    nf = F.shape[0]
    x = (V * stretch).reshape(-1)
    u = np.zeros(6 * nf)
    z = np.zeros(6 * nf)
    _, _, sums, dual, dcur, dprev = O.local_step(V, F, 100.0, x, u, z)
    area = O.mass_area(V, F)[1].sum()
    zs = 0.5 * (stretch + np.clip(stretch, 0.9, 1.1))
    np.testing.assert_allclose(sums[0], area * 2 * (stretch - zs) ** 2, rtol=1e-9, atol=1e-14)
    np.testing.assert_allclose(sums[1], area * 2 * stretch ** 2, rtol=1e-9)
    np.testing.assert_allclose(sums[2], area * 2 * zs ** 2, rtol=1e-9)


# ------------------------------------------------------------ broadphase


def test_candidate_filters_on_a_tiny_scene():
    """Two triangles of one cloth sharing an edge, a second cloth above in
    another group and a one-triangle obstacle in the first group, all boxes
    overlapping: the self pairs never share a vertex, pairs across groups do
    not exist, every point-triangle pair across objects of one group does
    (find_collision.cu:17-66,462-466), and pinned-only pairs are dropped."""
    Va = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float32)
    Fa = np.array([[0, 1, 2], [2, 1, 3]], np.int32)
    Vb = Va + np.float32([0.1, 0.1, 0.001])
    Vo = np.array([[0, 0, -0.001], [1, 0, -0.001], [0, 1, -0.001]], np.float32)
    Fo = np.array([[0, 1, 2]], np.int32)
    ident = np.arange(4, dtype=np.int32)
    offs = np.array([0, 4], np.int32)
    ow = O.OracleWorld({})
    ow.add_cloth(Va, Fa, [], ident, offs, coll=(1, 1, 0, 0.3, 0.3))
    ow.add_cloth(Vb, Fa, [], ident, offs, coll=(1, 1, 1, 0.3, 0.3))
    ow.add_obstacle(Vo, Fo, coll=(1, 1, 0, 0.3, 0.3))
    ow.finalize()
    pos = np.vstack([Va, Vb, Vo]).astype(np.float64).reshape(-1)
    pt, ee = ow.candidates(pos, pos)
    pts = {tuple(r) for r in pt.tolist()}
    # cloth a: vertices 0-3, cloth b (group 1): 4-7, obstacle: 8-10
    inb = lambda v: 4 <= v < 8  # noqa: E731
    assert all(inb(p) == all(inb(v) for v in t) for p, *t in pts)  # groups
    assert all(not ({p} & set(t)) for p, *t in pts)          # no shared vertex
    assert {(p, *t) for p, *t in pts if inb(p)} == {(7, 4, 5, 6), (4, 6, 5, 7)}
    self_pt = {(p, *t) for p, *t in pts if p < 4 and max(t) < 4}
    assert self_pt == {(3, 0, 1, 2), (0, 2, 1, 3)}
    cross = {(p, *t) for p, *t in pts if (p < 4) != (max(t) < 4)}
    assert cross == {(p, 8, 9, 10) for p in range(4)} | {
        (p, *f) for p in range(8, 11) for f in ([0, 1, 2], [2, 1, 3])}
    for a0, a1, b0, b1 in ee.tolist():
        assert inb(a0) == inb(b0)  # edges of cloth b only meet cloth b
        assert not (a0 >= 8 and b0 >= 8)  # no obstacle self pairs
        assert not ({a0, a1} & {b0, b1})
        assert (min(a0, a1), max(a0, a1)) < (min(b0, b1), max(b0, b1))
    # All of cloth a pinned: only its pairs with free vertices survive (none:
    # cloth a has no free vertex and the obstacle none either).
    ow2 = O.OracleWorld({})
    ow2.add_cloth(Va, Fa, [0, 1, 2, 3], ident, offs, coll=(1, 1, 0, 0.3, 0.3))
    ow2.add_obstacle(Vo, Fo, coll=(1, 1, 0, 0.3, 0.3))
    ow2.finalize()
    pt2, ee2 = ow2.candidates(np.vstack([Va, Vo]).astype(np.float64).reshape(-1),
                              np.vstack([Va, Vo]).astype(np.float64).reshape(-1))
    assert len(pt2) == 0 and len(ee2) == 0


# ------------------------------------------------------------ partition


@pytest.mark.parametrize("mesh", ["grid32", "grid80", "jitter", "sphere"])
def test_partition_invariants(mesh):
    """The product's MAS partition (graph_partition_test.cpp:11-30): parts are
    non-empty, hold at most 32 vertices, cover every vertex exactly once, and
    there are at least ceil(V / 32) of them."""
    if mesh == "grid32":
        V, F = scenes.grid_sheet(32)
    elif mesh == "grid80":
        V, F = scenes.grid_sheet(80)
    elif mesh == "jitter":
        V, F = scenes.grid_sheet(45, jitter=0.004, seed=9)
    else:
        V, F = scenes.icosphere(3, 0.3, (0, 0, 0))
    w = sp.World()
    assert w.init_context()  # host-side: the partition needs no device
    r, h = w.add_cloth(sp.ClothConfig(), sp.CollisionConfig(), V, F, [])
    assert r, r.to_string()
    r, perm, offs = w.get_partition(h, V.shape[0])
    assert r
    nv = V.shape[0]
    assert sorted(perm.tolist()) == list(range(nv))
    sizes = np.diff(offs)
    assert offs[0] == 0 and offs[-1] == nv
    assert sizes.min() >= 1 and sizes.max() <= 32
    assert len(sizes) >= -(-nv // 32)
