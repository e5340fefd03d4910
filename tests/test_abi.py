"""The C ABI library: every function declared in include/silk_hip.h is
exported, and the host-side contract (validation codes of
cuda_backend.cu:110-352, handles, caller-order readback before any device
work) holds without touching a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import silk_physics_amd as sp
from silk_physics_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "silk_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(silk_hip_\w+)\s*\(",
                                 text, re.M)))


def test_header_declares_the_abi():
    names = declared_functions()
    assert "silk_hip_step" in names and "silk_hip_create" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(sp.lib_path()))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_world_library_exports_world_api():
    path = os.path.join(os.path.dirname(str(sp.lib_path())), "libsilk.so")
    out = os.popen(f"nm -DC --defined-only {path}").read()
    for sym in ["silk::World::init", "silk::World::solver_step",
                "silk::World::add_cloth", "silk::World::get_cloth_position",
                "silk::World::set_global_config", "silk::World::add_obstacle",
                "silk::World::set_obstacle_position"]:
        assert sym in out, sym


@pytest.fixture
def ctx():
    lib = sp.load_library()
    h = C.c_void_p()
    assert lib.silk_hip_create(0, C.byref(h)) == 0
    yield lib, h
    lib.silk_hip_destroy(h)


def _add(lib, h, V, F, pins=None):
    pins = np.zeros(0, np.int32) if pins is None else np.asarray(pins, np.int32)
    handle = C.c_uint32(99)
    cp = sp.SilkHipClothParams(100.0, 1e-4, 0.1, 0.01)
    cc = sp.SilkHipCollisionParams(1, 1, 0, 0.3, 0.3)
    V = np.ascontiguousarray(V, np.float32).reshape(-1)
    F = np.ascontiguousarray(F, np.int32).reshape(-1)
    st = lib.silk_hip_add_cloth(
        h, C.byref(cp), C.byref(cc), V.ctypes.data_as(C.POINTER(C.c_float)),
        V.size // 3, F.ctypes.data_as(C.POINTER(C.c_int32)), F.size // 3,
        pins.ctypes.data_as(C.POINTER(C.c_int32)), pins.size, C.byref(handle))
    return st, handle.value


def test_add_cloth_validation_codes(ctx):
    lib, h = ctx
    V, F = scenes.grid_sheet(4)
    st, hd = _add(lib, h, V, F)
    assert st == 0 and hd != 0
    # index out of range -> InvalidMesh with handle 0
    bad = F.copy()
    bad[0, 0] = 10_000
    st, hd = _add(lib, h, V, bad)
    assert st == sp.ErrorCode.InvalidMesh + 1 and hd == 0
    # degenerate (repeated corner)
    bad = F.copy()
    bad[0, 1] = bad[0, 0]
    st, hd = _add(lib, h, V, bad)
    assert st == sp.ErrorCode.InvalidMesh + 1 and hd == 0
    # two disconnected components
    V2 = np.vstack([V, V + np.float32([5, 0, 0])])
    F2 = np.vstack([F, F + V.shape[0]])
    st, hd = _add(lib, h, V2, F2)
    assert st == sp.ErrorCode.InvalidMesh + 1 and hd == 0
    # non-finite vertex
    Vn = V.copy()
    Vn[3, 1] = np.nan
    st, hd = _add(lib, h, Vn, F)
    assert st == sp.ErrorCode.InvalidMesh + 1 and hd == 0
    # pin out of range: InvalidPin, handle left untouched as in
    # cuda_backend.cu:123-130
    st, hd = _add(lib, h, V, F, pins=[0, 1000])
    assert st == sp.ErrorCode.InvalidPin + 1 and hd == 99


def test_handles_and_readback_without_device_work(ctx):
    lib, h = ctx
    V, F = scenes.grid_sheet(3, jitter=0.01, seed=4)
    st, h1 = _add(lib, h, V, F)
    st2, h2 = _add(lib, h, V, F)
    assert st == 0 and st2 == 0 and h1 != h2 and h1 and h2
    out = np.zeros(V.size, np.float32)
    assert lib.silk_hip_get_positions(h, h1, out.ctypes.data_as(C.POINTER(C.c_float)),
                                      out.size) == 0
    np.testing.assert_array_equal(out, V.reshape(-1))
    # wrong size / handle
    assert lib.silk_hip_get_positions(h, h1, out.ctypes.data_as(C.POINTER(C.c_float)),
                                      out.size - 3) != 0
    assert lib.silk_hip_get_positions(h, 12345, out.ctypes.data_as(C.POINTER(C.c_float)),
                                      out.size) == sp.ErrorCode.InvalidHandle + 1
    assert lib.silk_hip_remove(h, h1, 1) == 0
    assert lib.silk_hip_remove(h, h1, 1) == sp.ErrorCode.InvalidHandle + 1
    # cloth handle is not an obstacle handle
    assert lib.silk_hip_remove(h, h2, 0) == sp.ErrorCode.InvalidHandle + 1


def test_step_without_device_reports_no_device(ctx):
    lib, h = ctx
    n = C.c_int32(0)
    lib.silk_hip_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a device is present")
    V, F = scenes.grid_sheet(3)
    _add(lib, h, V, F)
    assert lib.silk_hip_step(h, None) == 12  # SILK_HIP_ERR_NO_DEVICE


def test_world_api_host_contract():
    exe = os.path.join(os.path.dirname(str(sp.lib_path())), "world_api_test")
    out = os.popen(f"{exe} host").read()
    assert "FAIL" not in out, out
    assert out.count("PASS") >= 5, out


def test_set_global_rejects_unknown_solver_modes(ctx):
    """mas_precision, pcg_criterion, warm_start, ccd_zero_toi and admm_graph
    accept only their declared values; a learned warm start needs a loaded network."""
    lib, h = ctx
    ok = sp.GlobalConfig()._c()
    assert lib.silk_hip_set_global(h, C.byref(ok)) == 0
    for field, value in [("mas_precision", 3), ("mas_precision", -1),
                         ("pcg_criterion", 3), ("warm_start", 5), ("warm_projection", 2),
                         ("ccd_zero_toi", 2), ("ccd_zero_toi", -1),
                         ("admm_graph", 2), ("admm_graph", -1),
                         ("warm_start", sp.WARM_LEARNED)]:
        g = sp.GlobalConfig()._c()
        setattr(g, field, value)
        assert lib.silk_hip_set_global(h, C.byref(g)) == sp.ErrorCode.InvalidConfig + 1, field
        assert lib.silk_hip_last_error(h)
    for field, value in [("mas_precision", sp.MAS_FP16), ("mas_precision", sp.MAS_BF16),
                         ("pcg_criterion", sp.PCG_RHS), ("pcg_criterion", sp.PCG_TRUE),
                         ("warm_start", sp.WARM_EXTRAPOLATE),
                         ("warm_projection", sp.PROJECT_GALERKIN),
                         ("ccd_zero_toi", sp.CCD_ZERO_CONTACT),
                         ("admm_graph", sp.GRAPH_ON)]:
        g = sp.GlobalConfig()._c()
        setattr(g, field, value)
        assert lib.silk_hip_set_global(h, C.byref(g)) == 0, field
