"""Oracle CCD (restating interval_root_finder.cpp:391-510 and
ccd.cpp:139-257 in FP64) against closed forms, conservativeness by dense
sampling, and the single-precision instantiation."""
import numpy as np
import pytest

import oracle as O

MS = 1e-3
TOL = 1e-6


def err_for(p0, p1, vf):
    am = np.maximum(np.abs(p0).max(0), np.abs(p1).max(0))
    return O.ccd_filter(am, vf)


def test_vf_constant_velocity_closed_form():
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float64)
    p0 = np.vstack([[0.2, 0.2, 1.0], tri])
    p1 = np.vstack([[0.2, 0.2, -1.0], tri])
    hit, toi = O.ccd(True, p0, p1, MS, TOL, err_for(p0, p1, True), stack_cap=64)
    t_star = (1.0 - MS) / 2.0  # |z(t)| = |1 - 2t| reaches ms
    assert hit
    # conservative, and within the iteration-limited resolution of the
    # interval search
    assert t_star - 5e-4 <= toi <= t_star + 1e-12


def test_ee_crossing_closed_form():
    p0 = np.array([[-1, 0, 1.0], [1, 0, 1.0], [0, -1, 0], [0, 1, 0]], dtype=np.float64)
    p1 = p0.copy()
    p1[:2, 2] = -1.0
    hit, toi = O.ccd(False, p0, p1, MS, TOL, err_for(p0, p1, False), stack_cap=64)
    t_star = (1.0 - MS) / 2.0
    assert hit
    assert t_star - 5e-4 <= toi <= t_star + 1e-12


@pytest.mark.parametrize("vf", [True, False])
def test_no_hit_parallel_motion(vf):
    if vf:
        p0 = np.array([[0.2, 0.2, 1.0], [0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float64)
    else:
        p0 = np.array([[-1, 0, 1.0], [1, 0, 1.0], [0, -1, 0], [0, 1, 0]], dtype=np.float64)
    p1 = p0.copy()
    p1[:, 0] += 0.5  # everything slides along x together
    hit, _ = O.ccd(vf, p0, p1, MS, TOL, err_for(p0, p1, vf), stack_cap=64)
    assert not hit


def _min_linf_vf(p, t, n=41):
    x = (1 - t) * p[0] + t * p[1]  # 4x3 at time t
    g = np.linspace(0, 1, n)
    u, v = np.meshgrid(g, g, indexing="ij")
    m = (u + v) <= 1
    u, v = u[m], v[m]
    face = (1 - u - v)[:, None] * x[1] + u[:, None] * x[2] + v[:, None] * x[3]
    return np.abs(x[0] - face).max(1).min()


def point_triangle_distance(p, a, b, c):
    """Exact Euclidean distance (closest point on a triangle, Ericson 5.1.5)."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return np.linalg.norm(p - a)
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return np.linalg.norm(p - b)
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return np.linalg.norm(p - (a + d1 / (d1 - d3) * ab))
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return np.linalg.norm(p - c)
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return np.linalg.norm(p - (a + d2 / (d2 - d6) * ac))
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        w = (d4 - d3) / ((d4 - d3) + (d5 - d6))
        return np.linalg.norm(p - (b + w * (c - b)))
    den = 1.0 / (va + vb + vc)
    v, w = vb * den, vc * den
    return np.linalg.norm(p - (a + ab * v + ac * w))


def _random_hitting_vf(rng):
    """A vertex-face query that touches at some t* in (0.1, 0.9) and starts
    well outside the minimum separation (queries already inside it at t = 0
    report toi = 0, which the step ignores by design)."""
    while True:
        ts = rng.uniform(0.1, 0.9)
        x = rng.uniform(-1, 1, (4, 3))
        b = rng.dirichlet([1, 1, 1])
        x[0] = b[0] * x[1] + b[1] * x[2] + b[2] * x[3]
        v = rng.uniform(-0.5, 0.5, (4, 3))
        p0 = np.float32(x - ts * v).astype(np.float64)
        p1 = np.float32(x + (1 - ts) * v).astype(np.float64)
        if point_triangle_distance(*p0) > 4 * MS:
            return p0, p1


def test_vf_conservative_by_sampling():
    rng = np.random.default_rng(11)
    for _ in range(25):
        p0, p1 = _random_hitting_vf(rng)
        hit, toi = O.ccd(True, p0, p1, MS, TOL, err_for(p0, p1, True), stack_cap=64)
        assert hit
        # no sampled configuration before the reported TOI is inside ms
        for t in np.linspace(0.0, toi, 40)[:-1]:
            assert _min_linf_vf((p0, p1), t) > MS * (1 - 1e-9)


def test_bounded_pending_list_is_conservative():
    rng = np.random.default_rng(12)
    for _ in range(50):
        p0, p1 = _random_hitting_vf(rng)
        e = err_for(p0, p1, True)
        h_big, t_big = O.ccd(True, p0, p1, MS, TOL, e, stack_cap=1 << 20)
        h_64, t_64 = O.ccd(True, p0, p1, MS, TOL, e, stack_cap=64)
        assert h_big and h_64
        assert t_64 <= t_big


def test_f32_and_f64_agree_on_hits():
    rng = np.random.default_rng(13)
    for _ in range(50):
        p0, p1 = _random_hitting_vf(rng)
        e = err_for(p0, p1, True)
        h64, t64 = O.ccd(True, p0, p1, MS, TOL, e, stack_cap=1 << 20)
        h32, t32 = O.ccd_f32(True, p0, p1, MS, TOL, e)
        assert h64 and h32
        assert abs(t64 - t32) < 1e-4


def _vertical_vf(z0, z1):
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float64)
    return np.vstack([[0.2, 0.2, z0], tri]), np.vstack([[0.2, 0.2, z1], tri])


@pytest.mark.parametrize("zero_mode", [0, 1])
def test_zero_toi_refinement_closed_forms(zero_mode):
    """A point over a triangle's interior at height z0 moving straight down or
    up (its distance to the face is |z(t)|, z(t) = z0 + t (z1 - z0)).
    Starting within ms but beyond ms / 2, the ms / 2 search finds a time no
    later than |z| = ms / 2 in both modes.  Starting within ms / 2, the
    reference mode reports a root at t = 0 (no collision); the zero-separation
    search finds a time no later than the crossing z = 0, or no contact at
    all (toi = 1) when the point moves away.  TOIs are conservative (never
    after the closed form) and within the search's resolution of it."""
    def q(z0, z1):
        p0, p1 = _vertical_vf(z0, z1)
        return O.ccd(True, p0, p1, MS, TOL, err_for(p0, p1, True), stack_cap=64,
                     zero_mode=zero_mode, with_zero=True)

    hit, toi, zero = q(0.7 * MS, -0.7 * MS)
    t_half = (0.7 - 0.5) / 1.4
    assert hit and not zero and 0.7 * t_half <= toi <= t_half

    for z1, t_cross in ((-0.1 * MS, 0.5), (-1.0, 0.1 * MS / (1.0 + 0.1 * MS))):
        hit, toi, zero = q(0.1 * MS, z1)
        if zero_mode == 0:
            assert not hit and zero
        else:
            assert hit and not zero and 0.9 * t_cross <= toi <= t_cross

    hit, toi, zero = q(0.1 * MS, 1.0)
    if zero_mode == 0:
        assert not hit and zero
    else:
        assert hit and not zero and toi == 1.0
