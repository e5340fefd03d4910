"""Contact update by independent groups: when no contact couples two cloths,
the contacts are reordered by (group, colour) and one workgroup updates each
group (contacts of one colour share no movable vertex, so their order does
not matter).  The trajectory must equal the one-workgroup update bit for bit
after every step, on a batch of uneven sheets each over its own sphere."""
import numpy as np
import pytest

import silk_physics_amd as sp
from silk_physics_amd import scenes

from helpers import gpu_world

pytestmark = pytest.mark.gpu

KW = dict(acceleration_z=-9.81, dt=1 / 60, max_inner_iteration=10,
          linear_solver_rel_tol_min=1e-6, linear_solver_rel_tol_max=1e-6,
          initial_linear_rel_tol=1e-6, pcg_criterion=sp.PCG_TRUE)


def _world(sheets):
    w = gpu_world(**KW)
    hs = []
    for s in sheets:
        cc = sp.CollisionConfig(group=s["group"])
        r, h = w.add_cloth(sp.ClothConfig(), cc, s["V"], s["F"], s["pins"])
        assert r, r.to_string()
        r, _ = w.add_obstacle(cc, s["SV"], s["SF"])
        assert r, r.to_string()
        hs.append(h)
    return w, hs


def test_grouped_contact_update_matches_one_workgroup(monkeypatch):
    sheets = scenes.c4_scene(6, nmin=12, nmax=40)
    monkeypatch.setenv("SILK_HIP_CONTACT_GROUPS", "0")
    single, hs = _world(sheets)
    assert single.solver_step()  # the scene (and its switch) is built here
    monkeypatch.delenv("SILK_HIP_CONTACT_GROUPS")
    grouped, hg = _world(sheets)
    assert grouped.solver_step()
    groups = contacts = 0
    for k in range(12):
        for w in (single, grouped):
            r = w.solver_step()
            assert r, f"step {k}: {r.to_string()}"
        assert single.last_stats.contact_groups == 0
        groups = max(groups, grouped.last_stats.contact_groups)
        contacts += grouped.last_stats.contacts
        for s, a, b in zip(sheets, hs, hg):
            _, xa = single.get_cloth_position(a, n=s["V"].size)
            _, xb = grouped.get_cloth_position(b, n=s["V"].size)
            assert np.array_equal(xa, xb), f"step {k}: max diff {np.abs(xa - xb).max():.3e}"
    print(f"12 steps: up to {groups} contact groups, {contacts / 12:.0f} contacts per step")
    assert contacts > 0 and groups > 1
