"""The oracle's decision log and replay (the lockstep tests hold the GPU's
continuous state to the oracle given the GPU's discrete decisions): replaying
the oracle's own decisions reproduces its own step bit for bit, and
replaying the same TOIs without the contacts changes the step."""
import numpy as np

import oracle as O
from silk_physics_amd import scenes


def _world(n=10):
    V, F = scenes.grid_sheet(n, z=0.305)
    SV, SF = scenes.icosphere(1, 0.3, (0.5, 0.5, 0.0))
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    ow = O.OracleWorld(dict(az=f32(-9.81), dt=f32(1 / 60), max_inner=20,
                            ms_scale=f32(0.2), ccd_tol=f32(1e-6)))
    perm = np.arange(V.shape[0], dtype=np.int32)
    offs = np.arange(0, V.shape[0] + 1, 32, dtype=np.int32)
    if offs[-1] != V.shape[0]:
        offs = np.append(offs, V.shape[0]).astype(np.int32)
    ow.add_cloth(V, F, [], perm, offs)
    ow.add_obstacle(SV, SF)
    ow.finalize()
    return ow, V


def test_replaying_own_decisions_reproduces_the_step():
    ow, V = _world()
    x = V.astype(np.float64).reshape(-1)
    v = np.zeros_like(x)
    v[2::3] = -0.8
    seen = 0
    for _ in range(4):
        ow.set_state(0, x, v)
        rc, st = ow.step()
        assert rc == 0
        x1 = ow.positions(0)
        tois, contacts, _, admm, info = ow.decisions()
        seen += len(contacts)
        ow.set_state(0, x, v)
        ow.set_replay(tois, contacts, admm, info)
        rc, st2 = ow.step()
        assert rc == 0
        x2 = ow.positions(0)
        t2, c2, miss, a2, f2 = ow.decisions()
        assert miss == 0
        np.testing.assert_array_equal(t2, tois)
        np.testing.assert_array_equal(c2, contacts)
        np.testing.assert_array_equal(a2, admm)
        np.testing.assert_array_equal(f2, info)
        np.testing.assert_array_equal(x2, x1)
        v = (x1 - x) / np.float32(1 / 60)
        x = x1
    assert seen > 0


def test_replay_without_the_contacts_differs():
    ow, V = _world()
    x = V.astype(np.float64).reshape(-1)
    v = np.zeros_like(x)
    v[2::3] = -0.8
    ow.set_state(0, x, v)
    rc, _ = ow.step()
    x1 = ow.positions(0)
    tois, contacts, _, admm, info = ow.decisions()
    assert len(contacts) > 1
    ow.set_state(0, x, v)
    ow.set_replay(tois, contacts[:0])
    rc, _ = ow.step()
    assert rc == 0
    assert np.abs(ow.positions(0) - x1).max() > 0.0


def _fp32_log(tois):
    """The TOI log an FP32 solver writes for the same decisions: each TOI
    rounded to FP32, the final (accepted) one equal to that solver's FP32
    remaining interval (remaining -= 0.8 toi after each rejected TOI)."""
    t = [np.float32(v) for v in tois]
    rem = np.float32(1.0)
    for k in range(len(t) - 1):
        rem = np.float32(rem - np.float32(np.float32(0.8) * t[k]))
    t[-1] = rem
    return np.array(t, np.float64), rem


def test_replaying_fp32_tois_keeps_the_outer_iterations():
    """Replayed FP32 TOIs: the oracle takes the accept/reject decisions as the
    FP32 solver made them, so it runs exactly as many outer iterations and
    lands within FP32 rounding of its own step, including steps where the
    FP32 remaining interval is below the FP64 one (the final TOI would fail an
    FP64 acceptance test)."""
    ow, V = _world()
    x = V.astype(np.float64).reshape(-1)
    v = np.zeros_like(x)
    v[2::3] = -0.8
    below, multi = 0, 0
    for _ in range(12):
        ow.set_state(0, x, v)
        rc, st = ow.step()
        assert rc == 0
        x1 = ow.positions(0)
        tois, contacts, _, admm, info = ow.decisions()
        if len(tois) >= 2:
            multi += 1
            log, rem_f = _fp32_log(tois)
            rem_d = 1.0
            for t in tois[:-1]:
                rem_d -= 0.8 * t
            below += float(rem_f) < rem_d
            ow.set_state(0, x, v)
            ow.set_replay(log, contacts, admm)
            rc, st2 = ow.step()
            assert rc == 0
            assert st2["outer"] == st["outer"]
            assert len(ow.decisions()[0]) == len(tois)
            assert np.abs(ow.positions(0) - x1).max() < 1e-6
        v = (x1 - x) / np.float32(1 / 60)
        x = x1
    assert multi > 0 and below > 0, (multi, below)


def test_replayed_admm_counts_are_run_exactly():
    """The ADMM iteration count of each outer iteration is a decision of the
    convergence test: a replayed count is run as given (one more iteration
    than the test chose changes the step, the test's own count reproduces
    it)."""
    ow, V = _world()
    x = V.astype(np.float64).reshape(-1)
    v = np.zeros_like(x)
    v[2::3] = -0.8
    ow.set_state(0, x, v)
    rc, st = ow.step()
    x1 = ow.positions(0)
    tois, contacts, _, admm, info = ow.decisions()
    assert len(admm) == len(tois) and int(admm.sum()) == int(st["admm"])
    assert (admm < 20).any()  # the test stopped some outer iteration early
    more = admm.copy()
    more[np.argmin(admm)] += 1
    ow.set_state(0, x, v)
    ow.set_replay(tois, contacts, more)
    rc, st2 = ow.step()
    assert rc == 0
    np.testing.assert_array_equal(ow.decisions()[3], more)
    assert np.abs(ow.positions(0) - x1).max() > 0.0


def test_replayed_colours_and_masks_are_used():
    """A replayed step takes the contacts' colours and solved-vertex masks as
    given: a reversed colouring is reported back as replayed, empty masks
    (no vertex in any contact's solve) change the step, and the step's own
    info reproduces it bit for bit."""
    ow, V = _world()
    x = V.astype(np.float64).reshape(-1)
    v = np.zeros_like(x)
    v[2::3] = -0.8
    ow.set_state(0, x, v)
    rc, _ = ow.step()
    x1 = ow.positions(0)
    tois, contacts, _, admm, info = ow.decisions()
    assert len(info) == len(contacts) > 4 and (info & 15).any()
    own = info.copy()
    spread = (len(info) - np.arange(len(info), dtype=np.int32)) << 4 | (info & 15)
    ow.set_state(0, x, v)
    ow.set_replay(tois, contacts, admm, spread)
    rc, _ = ow.step()
    assert rc == 0
    np.testing.assert_array_equal(ow.decisions()[4], spread)
    off = info & ~15
    ow.set_state(0, x, v)
    ow.set_replay(tois, contacts, admm, off)
    rc, _ = ow.step()
    assert rc == 0
    np.testing.assert_array_equal(ow.decisions()[4], off)
    assert np.abs(ow.positions(0) - x1).max() > 0.0
    ow.set_state(0, x, v)
    ow.set_replay(tois, contacts, admm, own)
    rc, _ = ow.step()
    np.testing.assert_array_equal(ow.positions(0), x1)
