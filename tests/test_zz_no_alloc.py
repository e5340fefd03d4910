"""silk_hip_ccd_min_noalloc, the allocation-free variant of
silk_hip_debug_ccd_min, called through the C ABI: the same earliest TOI bit
for bit on two calls in a row (the second smaller, so stale scratch contents
would show), a query list with no hit (the minimum word is reset to 1 on
every call), and the capacity limit rejected with INVALID_CONFIG."""
import ctypes as C

import numpy as np
import pytest

import silk_physics_amd as sp
from silk_physics_amd import load_library

from test_gpu_collision import _queries
from test_oracle_ccd import MS, TOL

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 16384  # SILK_HIP_CCD_NOALLOC_MAX


def _noalloc(ctx, vf, p0, p1, err):
    lib = load_library()
    v = np.ascontiguousarray(vf, dtype=np.int32).reshape(-1)
    a = np.ascontiguousarray(p0, dtype=np.float64).reshape(-1)
    b = np.ascontiguousarray(p1, dtype=np.float64).reshape(-1)
    e = np.ascontiguousarray(err, dtype=np.float64).reshape(-1)
    out = C.c_double(-1.0)
    dp = C.POINTER(C.c_double)
    st = lib.silk_hip_ccd_min_noalloc(ctx, v.size, v.ctypes.data_as(C.POINTER(C.c_int32)),
                                      a.ctypes.data_as(dp), b.ctypes.data_as(dp),
                                      e.ctypes.data_as(dp), MS, TOL, 4096, C.byref(out))
    return st, out.value


def test_noalloc_matches_allocating_entry_point():
    w = sp.World()
    assert w.init(sp.Backend.GPU)
    vf, p0, p1, err = _queries(6000, seed=11)
    ref1 = w.debug_ccd_min(vf, p0, p1, err, MS, TOL)
    st, got1 = _noalloc(w._ctx, vf, p0, p1, err)
    assert st == 0
    print(f"6000 queries: allocating {ref1!r}, no-alloc {got1!r}")
    assert got1 == ref1 and ref1 < 1.0
    # Second call, smaller: the last 1500 queries in reverse order.
    k = 1500
    order = np.arange(vf.size - 1, vf.size - 1 - k, -1)
    sub = (vf[order], p0[order], p1[order], err[order])
    ref2 = w.debug_ccd_min(*sub, MS, TOL)
    st, got2 = _noalloc(w._ctx, *sub)
    assert st == 0
    print(f"{k} queries: allocating {ref2!r}, no-alloc {got2!r}")
    assert got2 == ref2
    # Queries without a hit: the minimum must be reset to 1, not kept.
    hit = np.array([w.debug_ccd_min(vf[i:i + 1], p0[i:i + 1], p1[i:i + 1], err[i:i + 1],
                                    MS, TOL) < 1.0 for i in range(200)])
    miss = np.flatnonzero(~hit)
    assert miss.size > 0
    st, got3 = _noalloc(w._ctx, vf[miss], p0[miss], p1[miss], err[miss])
    assert st == 0 and got3 == 1.0


def test_noalloc_rejects_more_than_capacity():
    w = sp.World()
    assert w.init(sp.Backend.GPU)
    n = NOALLOC_MAX + 1
    vf = np.ones(n, dtype=np.int32)
    p = np.zeros((n, 12))
    err = np.zeros((n, 3))
    st, _ = _noalloc(w._ctx, vf, p, p, err)
    assert st == sp.ErrorCode.InvalidConfig + 1
    assert b"capacity" in load_library().silk_hip_last_error(w._ctx)
