"""GPU collision path against the oracle: the FP64 CCD kernel bit for bit,
broadphase candidate sets exactly, a sheet falling on a sphere step by step
(lockstep from the same state) and free-running without penetration."""
import numpy as np
import pytest

import oracle as O
import silk_physics_amd as sp
from silk_physics_amd import scenes

from helpers import (coll_tuple, gpu_world, inside_convex, lockstep_step, oracle_for,
                     self_intersections, sheet_on_sphere)
from test_oracle_ccd import MS, TOL, err_for, point_triangle_distance

pytestmark = pytest.mark.gpu


def _f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def _queries(n, seed):
    rng = np.random.default_rng(seed)
    vf = rng.integers(0, 2, n).astype(np.int32)
    p0 = _f32(rng.uniform(-1, 1, (n, 12)))
    mv = _f32(rng.uniform(-1, 1, (n, 12)) * rng.choice([0.01, 0.3, 1.0], (n, 1)))
    p1 = _f32(p0 + mv)
    for i in range(0, n, 2):  # every other query touches at some t* in (0, 1)
        ts = rng.uniform(0.05, 0.95)
        x = rng.uniform(-1, 1, (4, 3))
        if vf[i]:
            b = rng.dirichlet([1, 1, 1])
            x[0] = b[0] * x[1] + b[1] * x[2] + b[2] * x[3]
        else:
            u = rng.uniform(0.1, 0.9)
            c = (1 - u) * x[0] + u * x[1]
            wv = rng.uniform(-0.5, 0.5, 3)
            x[2] = c + wv
            x[3] = c - rng.uniform(0.2, 2.0) * wv
        v = rng.uniform(-1, 1, (4, 3)) * rng.choice([0.05, 0.5])
        p0[i] = _f32((x - ts * v).reshape(-1))
        p1[i] = _f32((x + (1 - ts) * v).reshape(-1))
    err = np.array([err_for(p0[i].reshape(4, 3), p1[i].reshape(4, 3), bool(vf[i]))
                    for i in range(n)])
    return vf, p0, p1, err


def _near_queries(n, ms, seed):
    """Queries that start within ms of contact: a point at height h in
    (0.05, 0.95) ms over a triangle's interior, or an edge that far above
    another it crosses, then moving through, away or along (so that the
    refinements at ms / 2 and at zero separation decide them)."""
    rng = np.random.default_rng(seed)
    vf = (np.arange(n) % 2 == 0).astype(np.int32)
    p0 = np.zeros((n, 12))
    p1 = np.zeros((n, 12))
    for i in range(n):
        h = rng.uniform(0.05, 0.95) * ms
        if vf[i]:
            tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
            b = rng.dirichlet([2, 2, 2])
            x = np.vstack([b @ tri + [0, 0, h], tri])
        else:
            x = np.array([[-1, 0, h], [1, 0, h], [0, -1, 0], [0, 1, 0]], np.float64)
        x[:, :2] += rng.uniform(-1, 1, 2)
        dz = rng.choice([-1.0, 1.0, 0.0]) * rng.choice([0.2 * ms, 2 * ms, 0.5])
        v = np.zeros((4, 3))
        v[:2 if not vf[i] else 1, 2] = dz
        v[:, :2] += rng.uniform(-0.3, 0.3, 2) * rng.choice([0.0, 1.0])
        p0[i] = _f32(x.reshape(-1))
        p1[i] = _f32((x + v).reshape(-1))
    err = np.array([err_for(p0[i].reshape(4, 3), p1[i].reshape(4, 3), bool(vf[i]))
                    for i in range(n)])
    return vf, p0, p1, err


@pytest.mark.parametrize("zero_mode", [sp.CCD_ZERO_SKIP, sp.CCD_ZERO_CONTACT])
@pytest.mark.parametrize("ms", [MS, 1e-5])
def test_ccd_kernel_bit_exact_vs_oracle(ms, zero_mode):
    """The FP64 CCD kernel and the oracle, bit for bit, on random queries
    (half of them touching at some t in (0, 1)) and on queries that start
    within ms of contact, under both treatments of a root at t = 0."""
    w = gpu_world(ccd_zero_toi=zero_mode)
    a = _queries(3000, seed=int(ms * 1e6) + 1)
    b = _near_queries(1000, ms, seed=int(ms * 1e6) + 2)
    vf, p0, p1, err = (np.concatenate([x, y]) for x, y in zip(a, b))
    hit, toi = w.debug_ccd(vf, p0, p1, err, ms, TOL)
    ohit = np.zeros_like(hit)
    otoi = np.zeros_like(toi)
    zeros = 0
    for i in range(vf.size):
        h, t, z = O.ccd(bool(vf[i]), p0[i].reshape(4, 3), p1[i].reshape(4, 3), ms, TOL,
                        err[i], stack_cap=64, zero_mode=zero_mode, with_zero=True)
        ohit[i], otoi[i] = h, (t if h else 0.0)
        zeros += z
    near_hits = int(hit[3000:].sum())
    print(f"CCD ms={ms} zero mode {zero_mode}: {int(hit.sum())} hits of {vf.size} "
          f"({near_hits} of the near queries), oracle {int(ohit.sum())}; {zeros} "
          f"roots left at t = 0")
    assert ohit.sum() > vf.size // 4
    if zero_mode == sp.CCD_ZERO_SKIP:
        assert zeros > 100
    np.testing.assert_array_equal(hit, ohit)
    np.testing.assert_array_equal(toi, otoi)  # bit for bit


def _oracle_each(vf, p0, p1, err, ms):
    toi = np.zeros(vf.size)
    hit = np.zeros(vf.size, dtype=bool)
    for i in range(vf.size):
        h, t = O.ccd(bool(vf[i]), p0[i].reshape(4, 3), p1[i].reshape(4, 3), ms, TOL,
                     err[i], stack_cap=64)
        hit[i], toi[i] = h, (t if h else 0.0)
    return hit, toi


@pytest.mark.parametrize("order", ["as_drawn", "earliest_first", "earliest_last"])
def test_ccd_min_kernel_equals_oracle_minimum(order):
    """The step's earliest-TOI kernel cuts queries short once they cannot
    lower the running minimum; the minimum must still be the oracle's, bit
    for bit, whatever the order the queries come in."""
    w = gpu_world()
    vf, p0, p1, err = _queries(3000, seed=7)
    hit, toi = _oracle_each(vf, p0, p1, err, MS)
    pos = hit & (toi > 0)
    assert pos.sum() > 100
    key = np.where(pos, toi, 2.0)
    if order == "earliest_first":
        perm = np.argsort(key, kind="stable")
    elif order == "earliest_last":
        perm = np.argsort(-key, kind="stable")
    else:
        perm = np.arange(vf.size)
    expect = float(toi[pos].min())
    got = w.debug_ccd_min(vf[perm], p0[perm], p1[perm], err[perm], MS, TOL)
    assert got == expect, (got, expect)
    # Only non-colliding queries left: the minimum stays at 1.
    miss = ~hit
    assert w.debug_ccd_min(vf[miss], p0[miss], p1[miss], err[miss], MS, TOL) == 1.0


def test_ccd_kernel_closed_forms():
    w = gpu_world()
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float64)
    vf0 = np.vstack([[0.2, 0.2, 1.0], tri])
    vf1 = np.vstack([[0.2, 0.2, -1.0], tri])
    ee0 = np.array([[-1, 0, 1.0], [1, 0, 1.0], [0, -1, 0], [0, 1, 0]], dtype=np.float64)
    ee1 = ee0.copy()
    ee1[:2, 2] = -1.0
    miss0 = vf0.copy()
    miss1 = vf0.copy()
    miss1[:, 0] += 0.5
    p0 = np.stack([vf0, ee0, miss0]).reshape(3, 12)
    p1 = np.stack([vf1, ee1, miss1]).reshape(3, 12)
    is_vf = np.array([1, 0, 1], np.int32)
    err = np.array([err_for(p0[i].reshape(4, 3), p1[i].reshape(4, 3), bool(is_vf[i]))
                    for i in range(3)])
    hit, toi = w.debug_ccd(is_vf, p0, p1, err, MS, TOL)
    t_star = (1 - MS) / 2
    assert hit.tolist() == [1, 1, 0]
    assert t_star - 5e-4 <= toi[0] <= t_star and t_star - 5e-4 <= toi[1] <= t_star


def _pair_worlds(n=20, subdiv=2, z=0.32, global_kw=None):
    global_kw = global_kw or {}
    V, F, SV, SF = sheet_on_sphere(n, subdiv, z)
    w = gpu_world(**global_kw)
    cc = sp.CollisionConfig()
    r, h = w.add_cloth(sp.ClothConfig(), cc, V, F, np.zeros(0, np.int32))
    assert r, r.to_string()
    r, ho = w.add_obstacle(cc, SV, SF)
    assert r, r.to_string()
    r, perm, offs = w.get_partition(h, V.shape[0])
    ow = O.OracleWorld(oracle_for(global_kw))
    ow.add_cloth(V, F, [], perm, offs, coll=coll_tuple(cc))
    ow.add_obstacle(SV, SF, coll=coll_tuple(cc))
    ow.finalize()
    return V, F, SV, SF, w, h, ho, ow, perm


@pytest.mark.parametrize("trial", [0, 1, 2])
def test_broadphase_sets_equal_oracle(trial):
    V, F, SV, SF, w, h, ho, ow, perm = _pair_worlds()
    nvt = w.collider_vertex_count()
    assert nvt == V.shape[0] + SV.shape[0]
    rng = np.random.default_rng(100 + trial)
    base = np.zeros((nvt, 3))
    base[perm] = V
    base[V.shape[0]:] = SV
    a = _f32(base + rng.uniform(-0.02, 0.02, base.shape))
    b = _f32(a + rng.uniform(-0.03, 0.03, base.shape) * (trial + 1))
    gpt, gee = w.debug_candidates(a, b)
    opt, oee = ow.candidates(a.reshape(-1), b.reshape(-1))
    print(f"broadphase trial {trial}: PT {len(gpt)} EE {len(gee)}")
    assert len(opt) > 0 and len(oee) > 0
    assert sorted(map(tuple, gpt.tolist())) == sorted(map(tuple, opt.tolist()))
    assert sorted(map(tuple, gee.tolist())) == sorted(map(tuple, oee.tolist()))


@pytest.mark.parametrize("zero_mode", [sp.CCD_ZERO_SKIP, sp.CCD_ZERO_CONTACT])
def test_lockstep_steps_match_oracle(zero_mode):
    """Each step starts the oracle from the GPU's own state, so the error
    measured is that of one step of the collision path, not accumulated
    divergence of two chaotic trajectories.  The oracle steps twice: with
    its own discrete decisions (CCD TOIs, DCD contact set) and with the
    GPU's decisions replayed.  With the GPU's decisions every step is within
    1e-4 m; on its own decisions the oracle agrees to 1e-4 m on every step
    whose decisions match.  Where they differ, a threshold test (a distance
    against the activation distance, a barycentric sign, a dyadic TOI
    interval) fell the other way on inputs that differ by FP32 rounding; the
    decision functions themselves are held bit-exact on identical inputs
    (broadphase, CCD) or to FP32 bands (DCD) by the other tests here."""
    kw = dict(acceleration_z=-9.81, dt=1 / 60, max_inner_iteration=50,
              linear_solver_rel_tol_min=1e-6, linear_solver_rel_tol_max=1e-6,
              initial_linear_rel_tol=1e-6, admm_rel_tol=1e-5, admm_abs_tol=1e-7,
              ccd_zero_toi=zero_mode)
    V, F, SV, SF, w, h, ho, ow, perm = _pair_worlds(16, 2, 0.4, kw)
    dt = float(np.float32(1 / 60))
    x_prev = V.astype(np.float64).reshape(-1)
    v_prev = np.zeros_like(x_prev)
    rep, free_match, n_diff, contact_steps = [], [], 0, 0
    for i in range(40):
        xg, e_free, e_rep, same, missing = lockstep_step(w, h, ow, V.size, x_prev,
                                                          v_prev)
        contact_steps += w.last_stats.contacts > 0
        assert missing == 0, f"step {i}: {missing} replayed contacts not formed"
        rep.append(e_rep)
        if same:
            free_match.append(e_free)
        else:
            n_diff += 1
            print(f"step {i}: decisions differ, own-decision error {e_free:.3e}, "
                  f"replayed {e_rep:.3e}")
        v_prev = (xg - x_prev) / dt
        x_prev = xg
    rep = np.array(rep)
    print(f"lockstep, GPU decisions replayed: max {rep.max():.3e} median "
          f"{np.median(rep):.3e} over 40 steps; own decisions: {len(free_match)} "
          f"steps matched, max {max(free_match):.3e}; {n_diff} steps differ; "
          f"{contact_steps} steps with contacts")
    assert contact_steps >= 10
    assert rep.max() <= 1e-4
    assert len(free_match) >= 20 and max(free_match) <= 1e-4


def test_200_steps_with_replayed_decisions_within_1e4():
    """200 steps of a 10x10 sheet dropped on a sphere (its impact, sliding and
    rest), the oracle restarted from the GPU's state before every step and
    given the GPU's discrete decisions: every step's continuous result agrees
    within 1e-4 m.  An oracle that is never restarted drifts away from the
    GPU run once contacts are active (FP32 against FP64 rounding of about
    1e-7 m per step grows by orders of magnitude through contact and
    friction, and the replayed contacts then no longer describe the oracle's
    own geometry), so over many steps only the per-step error is a parity
    measure; the free fall without collision is held over 200 steps in
    test_gpu_pd.py."""
    kw = dict(acceleration_z=-9.81, dt=1 / 60, max_inner_iteration=20,
              linear_solver_rel_tol_min=1e-6, linear_solver_rel_tol_max=1e-6,
              initial_linear_rel_tol=1e-6)
    V, F, SV, SF, w, h, ho, ow, perm = _pair_worlds(10, 1, 0.36, kw)
    dt = float(np.float32(1 / 60))
    x_prev = V.astype(np.float64).reshape(-1)
    v_prev = np.zeros_like(x_prev)
    errs, contacts = [], 0
    for i in range(200):
        ow.set_state(0, x_prev, v_prev)
        assert w.solver_step()
        contacts += w.last_stats.contacts
        tois, cts, admm, info = w.decisions()
        ow.set_replay(tois, cts, admm, info)
        rc, _ = ow.step()
        assert rc == 0
        assert ow.decisions()[2] == 0
        _, xg = w.get_cloth_position(h, n=V.size)
        xg = xg.astype(np.float64)
        errs.append(float(np.abs(xg - ow.positions(0)).max()))
        v_prev = (xg - x_prev) / dt
        x_prev = xg
    errs = np.array(errs)
    print(f"200 steps, GPU decisions replayed from the GPU state: max error "
          f"{errs.max():.3e} (step {errs.argmax()}), median {np.median(errs):.3e}; "
          f"{contacts} contacts in all")
    assert contacts > 0
    assert errs.max() <= 1e-4


def test_free_running_sheet_does_not_penetrate_sphere():
    """60 free steps of a 16x16 sheet dropped on the sphere: after every step
    no cloth vertex is inside the (convex) obstacle and no cloth edge passes
    through a cloth triangle."""
    V, F, SV, SF = sheet_on_sphere(16, 2, 0.4)
    w = gpu_world(acceleration_z=-9.81, dt=1 / 60)
    cc = sp.CollisionConfig()
    r, h = w.add_cloth(sp.ClothConfig(), cc, V, F, np.zeros(0, np.int32))
    r, ho = w.add_obstacle(cc, SV, SF)
    touched = False
    closest = np.inf
    for i in range(60):
        assert w.solver_step()
        touched |= w.last_stats.contacts > 0
        _, x = w.get_cloth_position(h, n=V.size)
        s = inside_convex(x, SV, SF)
        closest = min(closest, s.min())
        assert s.min() > 0.0, (i, s.min())
        assert self_intersections(x, F) == 0, i
    print(f"60 steps: closest vertex {closest:.3e} m outside the sphere, no self-"
          f"intersection")
    assert touched


def test_coulomb_kernel_against_reference_kats_and_oracle():
    """The contact solve's Coulomb projection (solve_coulomb_contact,
    contact_constraints.cu:391-411) as the contact kernels run it, against
    the reference's known answers (friction_contact_test.cu:10-93, in
    tests/golden/reference_kats.json) and against the FP64 oracle on the
    reference's sweep grid and on random cases. FP32 tolerance: 1e-6
    absolute on the exact cases, 1e-5 x max(1, |u|) elsewhere (the
    projection is continuous across the cone boundary, so a boundary case
    decided the other way in FP32 still lands within it)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden",
                           "reference_kats.json")) as f:
        kat = json.load(f)
    w = gpu_world()
    ex = kat["coulomb_exact"]
    got = w.debug_coulomb([c["u"] for c in ex], [c["n"] for c in ex],
                          [c["mu"] for c in ex])
    for c, g in zip(ex, got):
        np.testing.assert_allclose(g, c["expected"], rtol=0, atol=1e-6, err_msg=c["name"])

    sw = kat["coulomb_sweep"]
    us = [u for _ in sw["mu"] for u in sw["u"]]
    mus = [m for m in sw["mu"] for _ in sw["u"]]
    rng = np.random.default_rng(7)
    ru = rng.uniform(-2, 2, (2000, 3)).astype(np.float32)
    rn = rng.normal(size=(2000, 3))
    rn = (rn / np.linalg.norm(rn, axis=1, keepdims=True)).astype(np.float32)
    rmu = rng.uniform(0, 2, 2000).astype(np.float32)
    u_all = np.vstack([np.asarray(us, dtype=np.float32), ru])
    n_all = np.vstack([np.tile(np.asarray(sw["n"], dtype=np.float32), (len(us), 1)), rn])
    mu_all = np.concatenate([np.asarray(mus, dtype=np.float32), rmu])
    got = w.debug_coulomb(u_all, n_all, mu_all)
    for i in range(len(mu_all)):
        u = u_all[i].astype(np.float64)
        want = O.coulomb(u, n_all[i].astype(np.float64), float(mu_all[i]))
        tol = 1e-5 * max(1.0, float(np.linalg.norm(u)))
        assert np.max(np.abs(got[i] - want)) <= tol, (i, u, n_all[i], mu_all[i], got[i], want)


@pytest.mark.parametrize("trial", [0, 1])
def test_dcd_records_equal_oracle_on_identical_inputs(trial):
    """DCD (dcd.cu:102-261) of the GPU and of the oracle on the same FP32
    positions: the same records, except pairs whose FP64 distance lies
    within an FP32 rounding band of the activation distance or whose FP64
    closest point lies that close to the triangle's edge (the PT test is
    FP32 on the GPU, as in the reference), and the same distances to FP32
    accuracy."""
    V, F, SV, SF, w, h, ho, ow, perm = _pair_worlds(16, 2, 0.301)
    nvt = w.collider_vertex_count()
    rng = np.random.default_rng(300 + trial)
    base = np.zeros((nvt, 3))
    base[perm] = V
    base[V.shape[0]:] = SV
    # The sheet laid over the sphere's cap, each vertex 0.2-3.5 mm above the
    # sphere (r = 0.3), so that many PT and EE pairs fall on either side of
    # the activation distance.
    pos = base.copy()
    nv = V.shape[0]
    r = np.hypot(pos[:nv, 0] - 0.5, pos[:nv, 1] - 0.5)
    pos[:nv, 2] = np.where(r < 0.27, np.sqrt(np.maximum(0.09 - r * r, 0.0)), 0.3)
    pos[:nv, 2] += rng.uniform(2e-4, 3.5e-3, nv)
    pos = _f32(pos)
    dhat = float(np.float32(0.05 / 16 * 1.02))
    grec, gdist = w.debug_dcd(pos, pos, 0.0, dhat)
    orec, odist = ow.dcd(pos.reshape(-1).astype(np.float64),
                         pos.reshape(-1).astype(np.float64), 0.0, dhat)

    def key(r):
        t, a, b, c, d = map(int, r)
        if t == 0:
            return (t, a, *sorted((b, c, d)))
        e1, e2 = tuple(sorted((a, b))), tuple(sorted((c, d)))
        return (t, *min(e1, e2), *max(e1, e2))

    g = {key(r): float(d) for r, d in zip(grec, gdist)}
    o = {key(r): float(d) for r, d in zip(orec, odist)}
    both = set(g) & set(o)
    n_pt = sum(1 for k in both if k[0] == 0)
    print(f"DCD trial {trial}: GPU {len(g)} records, oracle {len(o)}, common {len(both)} "
          f"({n_pt} PT)")
    assert len(both) > 200 and n_pt > 20
    dd = max(abs(g[k] - o[k]) for k in both)
    assert dd <= 2e-6, dd
    band = 1e-4 * dhat
    for k in set(g) ^ set(o):
        d = g.get(k, o.get(k))
        # Only near-threshold pairs may differ: the distance within the band
        # of dhat or within FP32 rounding of zero (a pair at distance 0 has
        # no normal and is dropped, dcd.cu:39,120), or a PT pair whose
        # projection sits on the triangle edge (checked on the oracle's FP64
        # geometry).
        near_dhat = abs(d - dhat) <= band or d <= 1e-7
        near_edge = False
        if k[0] == 0:
            x = pos.reshape(-1, 3).astype(np.float64)
            p, a, b, c = x[k[1]], x[k[2]], x[k[3]], x[k[4]]
            m = np.array([b - a, c - a]).T
            uv, *_ = np.linalg.lstsq(m, p - a, rcond=None)
            near_edge = min(uv[0], uv[1], 1 - uv[0] - uv[1]) > -1e-5 and \
                min(abs(uv[0]), abs(uv[1]), abs(1 - uv[0] - uv[1])) <= 1e-5
        assert near_dhat or near_edge, (k, d, dhat)


def test_device_bvh_reordering_keeps_the_trajectory(monkeypatch):
    """The per-step Morton re-ordering of the trees' primitives and query
    orders on the device (segmented by object, stable) changes only the order
    in which the broadphase visits pairs: 40 steps of a 16x16 sheet falling
    on the sphere, with the re-ordering and with the set-up order kept,
    give bit-identical positions (candidate sets, sorted contacts and the
    minimum TOI do not depend on the visiting order)."""
    def run(rebuild):
        monkeypatch.setenv("SILK_HIP_BVH_REBUILD", "1" if rebuild else "0")
        V, F, SV, SF = sheet_on_sphere(16, 2, 0.36)
        w = gpu_world(acceleration_z=-9.81, dt=1 / 60, max_inner_iteration=20)
        cc = sp.CollisionConfig()
        r, h = w.add_cloth(sp.ClothConfig(), cc, V, F, np.zeros(0, np.int32))
        assert r
        r, _ = w.add_obstacle(cc, SV, SF)
        assert r
        contacts = 0
        for _ in range(40):
            assert w.solver_step()
            contacts += w.last_stats.contacts
        return w.get_cloth_position(h, n=V.size)[1], contacts

    x1, c1 = run(True)
    x0, c0 = run(False)
    print(f"re-ordered vs set-up order: {c1} / {c0} contacts over 40 steps")
    assert c1 > 0 and c1 == c0
    np.testing.assert_array_equal(x1, x0)


def _first_fit(recs: np.ndarray, info: np.ndarray) -> np.ndarray:
    """Sequential first-fit colouring of a step's contacts in contact order:
    a contact takes the smallest colour >= 1 not held by an earlier contact
    that shares one of its solved vertices (mask bits of info)."""
    held: dict[int, set] = {}
    out = np.zeros(len(recs), np.int64)
    for c in range(len(recs)):
        vs = [int(recs[c, 2 + k]) for k in range(4) if info[c] >> k & 1]
        used = set().union(*(held.get(v, set()) for v in vs)) if vs else set()
        col = 1
        while col in used:
            col += 1
        out[c] = col
        for v in vs:
            held.setdefault(v, set()).add(col)
    return out


@pytest.mark.parametrize("where", ["lds", "global"])
def test_device_colouring_is_first_fit_in_contact_order(where, monkeypatch):
    """The device colouring (contact_coloring.cu:69-373 replaced by
    first-fit Jones-Plassmann rounds with the contact index as priority) equals
    the sequential first-fit colouring of the step's contacts in contact
    order, call after call (earlier contacts keep their colours), so no two
    contacts sharing a solved vertex share a colour; with the vertex masks in
    LDS and in global memory (SILK_HIP_COLOR_GLOBAL)."""
    if where == "global":
        monkeypatch.setenv("SILK_HIP_COLOR_GLOBAL", "1")
    s = scenes.c2_scene(40, z=0.305)
    w = gpu_world(acceleration_z=-9.81, dt=1 / 60, max_inner_iteration=10)
    cc = sp.CollisionConfig()
    r, h = w.add_cloth(sp.ClothConfig(), cc, s["V"], s["F"], s["pins"])
    assert r, r.to_string()
    r, _ = w.add_obstacle(cc, s["SV"], s["SF"])
    assert r, r.to_string()
    total = calls = max_col = 0
    for i in range(24):
        r = w.solver_step()
        assert r, f"step {i}: {r.to_string()}"
        _, recs, _, info = w.decisions()
        if len(recs) == 0:
            continue
        col = info >> 4
        want = _first_fit(recs, info & 15)
        assert np.array_equal(col, want), f"step {i}: {np.flatnonzero(col != want)[:8]}"
        assert w.last_stats.contact_colors == col.max(), f"step {i}"
        total += len(recs)
        calls += len(np.unique(recs[:, 0]))
        max_col = max(max_col, int(col.max()))
    print(f"device colouring ({where}): {total} contacts over {calls} DCD calls, up to {max_col} "
          f"colours, all first-fit in contact order")
    assert total > 1000 and calls > 24
