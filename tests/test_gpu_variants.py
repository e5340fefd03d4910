"""Every code path a large scene can take, forced on the 16x16 lockstep
scene: the four contact Gauss-Seidel kernels (contacts and corrections in
registers with the normals in LDS or in registers, positions in LDS, or
everything in global memory; SILK_HIP_CONTACT_VARIANT) and the multi-launch
PCG that replaces the persistent one when its grid does not fit the device
(SILK_HIP_PCG=launch).  Each is held to the oracle replaying its decisions."""
import os

import numpy as np
import pytest

from helpers import lockstep_step
from test_gpu_collision import _pair_worlds

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("env,value", [("SILK_HIP_CONTACT_VARIANT", "0"),
                                       ("SILK_HIP_CONTACT_VARIANT", "1"),
                                       ("SILK_HIP_CONTACT_VARIANT", "2"),
                                       ("SILK_HIP_CONTACT_VARIANT", "3"),
                                       ("SILK_HIP_PCG", "launch")])
def test_forced_path_matches_oracle(env, value, monkeypatch):
    monkeypatch.setenv(env, value)
    kw = dict(acceleration_z=-9.81, dt=1 / 60, max_inner_iteration=30,
              linear_solver_rel_tol_min=1e-6, linear_solver_rel_tol_max=1e-6,
              initial_linear_rel_tol=1e-6)
    V, F, SV, SF, w, h, ho, ow, perm = _pair_worlds(16, 2, 0.33, kw)
    if env == "SILK_HIP_PCG":
        assert w.scene_info()["pcg_workgroups"] == 0
    dt = float(np.float32(1 / 60))
    x_prev = V.astype(np.float64).reshape(-1)
    v_prev = np.zeros_like(x_prev)
    rep, contacts = [], 0
    for i in range(12):
        xg, _, e_rep, _, missing = lockstep_step(w, h, ow, V.size, x_prev, v_prev)
        assert missing == 0
        contacts += w.last_stats.contacts
        rep.append(e_rep)
        v_prev = (xg - x_prev) / dt
        x_prev = xg
    print(f"{env}={value}: replayed lockstep max {max(rep):.3e}, {contacts} contacts")
    assert contacts > 0
    assert max(rep) <= 1e-4
