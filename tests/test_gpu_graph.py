"""The device ADMM loop captured into a hipGraph (admm_graph = GRAPH_ON).

Every outer iteration's loop is captured anew and the scene's executable
graph is updated in place, so each graph launch replays kernels whose
arguments (contact counts and colour ranges, the parity of the z/u buffers)
changed since the last one, with the step's own eager work (CCD, DCD,
contact assembly and colouring, host synchronisations) between replays.
The trajectory must equal the eagerly enqueued one bit for bit after every
step, on a colliding sheet and on a batch of uneven sheets in one launch."""
import numpy as np
import pytest

import silk_physics_amd as sp
from silk_physics_amd import scenes

from helpers import gpu_world

pytestmark = pytest.mark.gpu

KW = dict(acceleration_z=-9.81, dt=1 / 60, max_inner_iteration=10,
          linear_solver_rel_tol_min=1e-6, linear_solver_rel_tol_max=1e-6,
          initial_linear_rel_tol=1e-6, pcg_criterion=sp.PCG_TRUE)


def _worlds(sheets, **kw):
    out = []
    for graph in (sp.GRAPH_OFF, sp.GRAPH_ON):
        w = gpu_world(admm_graph=graph, **kw)
        hs = []
        for s in sheets:
            cc = sp.CollisionConfig(group=s.get("group", 0))
            r, h = w.add_cloth(sp.ClothConfig(), cc, s["V"], s["F"], s["pins"])
            assert r, r.to_string()
            r, _ = w.add_obstacle(cc, s["SV"], s["SF"])
            assert r, r.to_string()
            hs.append(h)
        out.append((w, hs))
    return out


def _run(sheets, steps, **kw):
    (eager, he), (graph, hg) = _worlds(sheets, **kw)
    launches = inst = outer = contacts = 0
    for k in range(steps):
        for w in (eager, graph):
            r = w.solver_step()
            assert r, f"step {k}: {r.to_string()}"
        st = graph.last_stats
        assert eager.last_stats.graph_launches == 0
        # One graph launch per outer iteration (the ADMM loop of each).
        assert st.graph_launches == st.outer_iterations, k
        assert st.pcg_unverified_solves == 0
        launches += st.graph_launches
        inst += st.graph_instantiations
        outer += st.outer_iterations
        contacts += st.contacts
        assert st.pcg_iterations == eager.last_stats.pcg_iterations, k
        for s, a, b in zip(sheets, he, hg):
            _, xe = eager.get_cloth_position(a, n=s["V"].size)
            _, xg = graph.get_cloth_position(b, n=s["V"].size)
            assert np.array_equal(xe, xg), f"step {k}: max diff {np.abs(xe - xg).max():.3e}"
    print(f"{steps} steps: {launches} graph launches, {inst} instantiations, "
          f"{outer} outer iterations, {contacts / steps:.0f} contacts per step")
    return launches, inst, contacts


def test_graph_matches_eager_colliding_sheet():
    s = scenes.c2_scene(24, z=0.31)
    launches, inst, contacts = _run([s], 24, **KW)
    assert contacts > 0
    # The executable graph is updated in place, not rebuilt, on most launches.
    assert inst < launches


def test_graph_matches_eager_batched_sheets():
    sheets = scenes.c4_scene(6, nmin=12, nmax=40)
    launches, inst, contacts = _run(sheets, 10, **KW)
    assert contacts > 0
    assert inst < launches


def test_graph_off_while_timing():
    s = scenes.c2_scene(16, z=0.33)
    (_, _), (w, _) = _worlds([s], **KW)
    w.set_timing(True)
    assert w.solver_step()
    assert w.last_stats.graph_launches == 0
    w.set_timing(False)
    assert w.solver_step()
    assert w.last_stats.graph_launches == w.last_stats.outer_iterations
