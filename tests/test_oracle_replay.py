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
        tois, contacts, _ = ow.decisions()
        seen += len(contacts)
        ow.set_state(0, x, v)
        ow.set_replay(tois, contacts)
        rc, st2 = ow.step()
        assert rc == 0
        x2 = ow.positions(0)
        t2, c2, miss = ow.decisions()
        assert miss == 0
        np.testing.assert_array_equal(t2, tois)
        np.testing.assert_array_equal(c2, contacts)
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
    tois, contacts, _ = ow.decisions()
    assert len(contacts) > 1
    ow.set_state(0, x, v)
    ow.set_replay(tois, contacts[:0])
    rc, _ = ow.step()
    assert rc == 0
    assert np.abs(ow.positions(0) - x1).max() > 0.0
