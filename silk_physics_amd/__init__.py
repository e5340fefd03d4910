"""Python mirror of the silk::World API over the C ABI (include/silk_hip.h).

The classes and method names follow silk/include/silk/silk.hpp (World,
GlobalConfig, ClothConfig, CollisionConfig, Result, ErrorCode); calls never
raise on validation errors, they return a Result, as the reference does.

The native library is loaded from ``silk_physics_amd/lib``; it must have been
built with ``silk_physics_amd.build.build()`` (or ``__graft_entry__.build()``).
There is no fallback implementation: without the library every World call
raises ImportError.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

_LIB_DIR = Path(__file__).resolve().parent / "lib"
_lib = None


def lib_path() -> Path:
    return _LIB_DIR / "libsilk_hip.so"


def load_library():
    """Load libsilk_hip.so (once) and declare the C ABI signatures."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not path.exists():
        raise ImportError(f"{path} not built; run __graft_entry__.build()")
    lib = C.CDLL(str(path))
    i32, i64, u32 = C.c_int32, C.c_int64, C.c_uint32
    vp = C.c_void_p
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int32)
    dp = C.POINTER(C.c_double)
    sig = {
        "silk_hip_device_count": (C.c_int, [ip]),
        "silk_hip_create": (C.c_int, [i32, C.POINTER(vp)]),
        "silk_hip_destroy": (C.c_int, [vp]),
        "silk_hip_last_error": (C.c_char_p, [vp]),
        "silk_hip_set_global": (C.c_int, [vp, C.POINTER(SilkHipGlobal)]),
        "silk_hip_add_cloth": (C.c_int, [vp, C.POINTER(SilkHipClothParams),
                                         C.POINTER(SilkHipCollisionParams),
                                         fp, i64, ip, i64, ip, i64,
                                         C.POINTER(u32)]),
        "silk_hip_add_obstacle": (C.c_int, [vp, C.POINTER(SilkHipCollisionParams),
                                            fp, i64, ip, i64, C.POINTER(u32)]),
        "silk_hip_remove": (C.c_int, [vp, u32, i32]),
        "silk_hip_set_cloth_params": (C.c_int, [vp, u32,
                                                C.POINTER(SilkHipClothParams)]),
        "silk_hip_set_collision_params": (C.c_int, [vp, u32, i32,
                                                    C.POINTER(SilkHipCollisionParams)]),
        "silk_hip_set_pin_index": (C.c_int, [vp, u32, ip, i64]),
        "silk_hip_set_pin_positions": (C.c_int, [vp, u32, fp, i64]),
        "silk_hip_set_obstacle_positions": (C.c_int, [vp, u32, fp, i64]),
        "silk_hip_step": (C.c_int, [vp, C.POINTER(SilkHipStepStats)]),
        "silk_hip_reset": (C.c_int, [vp]),
        "silk_hip_get_positions": (C.c_int, [vp, u32, fp, i64]),
        "silk_hip_get_partition": (C.c_int, [vp, u32, ip, i64, ip, i64, ip]),
        "silk_hip_get_timing": (C.c_int, [vp, i32, C.POINTER(C.c_double),
                                          C.POINTER(C.c_int64), i32]),
        "silk_hip_collider_vertex_count": (C.c_int, [vp, C.POINTER(i64)]),
        "silk_hip_set_timing": (C.c_int, [vp, i32]),
        "silk_hip_get_scene_info": (C.c_int, [vp, C.POINTER(i64), i32]),
        "silk_hip_debug_candidates": (C.c_int, [vp, fp, fp, i64, ip, i64,
                                                C.POINTER(i64), ip, i64,
                                                C.POINTER(i64)]),
        "silk_hip_debug_dcd": (C.c_int, [vp, fp, fp, i64, C.c_float, C.c_float, ip,
                                         fp, i64, C.POINTER(i64)]),
        "silk_hip_debug_decisions": (C.c_int, [vp, fp, ip, i32, ip, ip, ip, i64,
                                               C.POINTER(i64)]),
        "silk_hip_debug_ccd": (C.c_int, [vp, i64, ip, dp, dp, dp, C.c_double,
                                         C.c_double, i32, ip, dp]),
        "silk_hip_debug_coulomb": (C.c_int, [vp, i64, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_void_p]),
        "silk_hip_debug_ccd_min": (C.c_int, [vp, i64, ip, dp, dp, dp, C.c_double,
                                             C.c_double, i32, dp]),
        "silk_hip_debug_pcg": (C.c_int, [vp, fp, i64, fp, i64, C.c_float, fp, ip]),
        "silk_hip_debug_copy": (C.c_int, [vp, i64, i32, dp]),
        "silk_hip_ccd_min_noalloc": (C.c_int, [vp, i64, ip, dp, dp, dp, C.c_double,
                                               C.c_double, i32, dp]),
        "silk_hip_set_warm_start_net": (C.c_int, [vp, fp, i64]),
        "silk_hip_autotune": (C.c_int, [vp, ip, C.c_int32]),
        "silk_hip_get_warm_start_net": (C.c_int, [vp, fp, i64]),
        "silk_hip_warm_start_record": (C.c_int, [vp, i64, C.c_float]),
        "silk_hip_warm_start_samples": (C.c_int, [vp, C.POINTER(i64)]),
        "silk_hip_warm_start_train": (C.c_int, [vp, i32, i32, C.c_float, C.c_uint64, fp]),
        "silk_hip_debug_warm_net": (C.c_int, [vp, i64, fp, fp, fp, fp, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


# ---------------------------------------------------------------- C structs

# Values of GlobalConfig.mas_precision, .pcg_criterion and .warm_start
# (include/silk_hip.h).
MAS_FP32, MAS_FP16, MAS_BF16 = 0, 1, 2
PCG_RR0, PCG_RHS, PCG_TRUE = 0, 1, 2
WARM_PREVIOUS, WARM_LEARNED, WARM_EXTRAPOLATE = 0, 1, 2
PROJECT_RESIDUAL, PROJECT_GALERKIN = 0, 1
CCD_ZERO_SKIP, CCD_ZERO_CONTACT = 0, 1
GRAPH_OFF, GRAPH_ON = 0, 1
DEVICE_CPU = -1  # silk_hip_create device of the CPU backend
WARM_NET_PARAMS = 5313  # SILK_HIP_WARM_NET_PARAMS
AUTOTUNE_REPORT = 10  # SILK_HIP_AUTOTUNE_REPORT

class SilkHipClothParams(C.Structure):
    _fields_ = [("elastic_stiffness", C.c_float), ("bending_stiffness", C.c_float),
                ("density", C.c_float), ("damping", C.c_float)]


class SilkHipCollisionParams(C.Structure):
    _fields_ = [("is_collision_on", C.c_int32), ("is_self_collision_on", C.c_int32),
                ("group", C.c_int32), ("restitution", C.c_float),
                ("friction", C.c_float)]


class SilkHipGlobal(C.Structure):
    _fields_ = [("acceleration", C.c_float * 3),
                ("max_outer_iteration", C.c_int32), ("max_inner_iteration", C.c_int32),
                ("dt", C.c_float),
                ("linear_solver_abs_tol", C.c_float),
                ("linear_solver_rel_tol_min", C.c_float),
                ("linear_solver_rel_tol_max", C.c_float),
                ("initial_linear_rel_tol", C.c_float),
                ("admm_abs_tol", C.c_float), ("admm_rel_tol", C.c_float),
                ("pcg_max_iter", C.c_int32), ("contact_sweeps", C.c_int32),
                ("ccd_max_iter", C.c_int32), ("ccd_tolerance", C.c_float),
                ("ccd_min_separation_scale", C.c_float),
                ("dcd_activation_scale", C.c_float),
                ("mas_precision", C.c_int32), ("true_residual_stats", C.c_int32),
                ("pcg_criterion", C.c_int32), ("warm_start", C.c_int32),
                ("warm_projection", C.c_int32), ("ccd_zero_toi", C.c_int32),
                ("admm_graph", C.c_int32), ("reserved", C.c_int32)]


class SilkHipStepStats(C.Structure):
    _fields_ = [("outer_iterations", C.c_int32), ("admm_iterations", C.c_int32),
                ("pcg_iterations", C.c_int32), ("pcg_solves", C.c_int32),
                ("contacts", C.c_int32), ("contact_colors", C.c_int32),
                ("ccd_candidates", C.c_int32), ("dcd_candidates", C.c_int32),
                ("min_toi", C.c_float), ("last_pcg_rel_residual", C.c_float),
                ("max_true_rel_residual", C.c_float),
                ("max_committed_rel_residual", C.c_float),
                ("pcg_max_solve_iterations", C.c_int32),
                ("pcg_stalled_solves", C.c_int32), ("warm_start_kept", C.c_int32),
                ("warm_start_solves", C.c_int32), ("pcg_restarts", C.c_int32),
                ("pcg_unverified_solves", C.c_int32), ("ccd_zero_toi_pairs", C.c_int32),
                ("graph_launches", C.c_int32), ("graph_instantiations", C.c_int32),
                ("contact_groups", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


# ------------------------------------------------------- API mirror (silk.hpp)

class ErrorCode(enum.IntEnum):
    Unknown = 0
    NotInitialized = 1
    InvalidConfig = 2
    TooManyBody = 3
    InvalidHandle = 4
    InvalidMesh = 5
    InvalidPin = 6
    InvalidPosition = 7
    CholeskyDecompositionFail = 8
    NeedInitSolverFirst = 9
    IterativeSolveFail = 10
    NoCudaSupport = 11


@dataclass
class Result:
    success: bool = True
    code: ErrorCode = ErrorCode.Unknown
    detail: str = ""

    def __bool__(self) -> bool:
        return self.success

    def to_string(self) -> str:
        if self.success:
            return "OK"
        if not self.detail:
            return self.code.name
        return f"{self.code.name}. Reason: {self.detail}."


class Backend(enum.IntEnum):
    CPU = 0
    GPU = 1


@dataclass
class CollisionConfig:
    is_collision_on: bool = True
    is_self_collision_on: bool = True
    group: int = 0
    restitution: float = 0.3
    friction: float = 0.3

    def _c(self) -> SilkHipCollisionParams:
        return SilkHipCollisionParams(int(self.is_collision_on),
                                      int(self.is_self_collision_on), self.group,
                                      self.restitution, self.friction)


@dataclass
class ClothConfig:
    elastic_stiffness: float = 100.0
    bending_stiffness: float = 0.0001
    density: float = 0.1
    damping: float = 0.01

    def _c(self) -> SilkHipClothParams:
        return SilkHipClothParams(self.elastic_stiffness, self.bending_stiffness,
                                  self.density, self.damping)


@dataclass
class GlobalConfig:
    acceleration_x: float = 0.0
    acceleration_y: float = 0.0
    acceleration_z: float = -10.0
    max_outer_iteration: int = 100
    max_inner_iteration: int = 100
    dt: float = 1.0 / 60.0
    linear_solver_abs_tol: float = 1e-7
    linear_solver_rel_tol_min: float = 1e-6
    linear_solver_rel_tol_max: float = 1e-3
    initial_linear_rel_tol: float = 1e-2
    admm_abs_tol: float = 1e-5
    admm_rel_tol: float = 1e-3
    # Extensions of this library (0 = reference default).
    pcg_max_iter: int = 0
    contact_sweeps: int = 0
    ccd_max_iter: int = 0
    ccd_tolerance: float = 0.0
    ccd_min_separation_scale: float = 0.0
    dcd_activation_scale: float = 0.0
    mas_precision: int = 0        # MAS_FP32 / MAS_FP16 / MAS_BF16
    true_residual_stats: int = 0
    pcg_criterion: int = 0        # PCG_RR0 (reference) / PCG_RHS / PCG_TRUE
    warm_start: int = 0           # WARM_PREVIOUS (reference) / WARM_LEARNED /
                                  # WARM_EXTRAPOLATE
    warm_projection: int = 0      # PROJECT_RESIDUAL / PROJECT_GALERKIN
    ccd_zero_toi: int = 0         # CCD_ZERO_SKIP (reference) / CCD_ZERO_CONTACT
    admm_graph: int = 0           # GRAPH_OFF / GRAPH_ON (ADMM loop as a hipGraph)

    def _c(self) -> SilkHipGlobal:
        g = SilkHipGlobal()
        g.acceleration[0] = self.acceleration_x
        g.acceleration[1] = self.acceleration_y
        g.acceleration[2] = self.acceleration_z
        for name in ("max_outer_iteration", "max_inner_iteration", "dt",
                     "linear_solver_abs_tol", "linear_solver_rel_tol_min",
                     "linear_solver_rel_tol_max", "initial_linear_rel_tol",
                     "admm_abs_tol", "admm_rel_tol", "pcg_max_iter",
                     "contact_sweeps", "ccd_max_iter", "ccd_tolerance",
                     "ccd_min_separation_scale", "dcd_activation_scale",
                     "mas_precision", "true_residual_stats", "pcg_criterion",
                     "warm_start", "warm_projection", "ccd_zero_toi", "admm_graph"):
            setattr(g, name, getattr(self, name))
        return g


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class World:
    """silk::World (silk/include/silk/silk.hpp:83-239) over the C ABI."""

    def __init__(self) -> None:
        self._ctx = None
        self.last_stats = SilkHipStepStats()

    def __del__(self) -> None:
        self.clear()

    # -- global
    def init(self, backend: Backend = Backend.GPU) -> Result:
        lib = load_library()
        if backend == Backend.CPU:
            # The reference's CPU solver on the host's cores (same C ABI).
            return self.init_context(DEVICE_CPU)
        n = C.c_int32(0)
        lib.silk_hip_device_count(C.byref(n))
        if n.value <= 0:
            return Result(False, ErrorCode.NoCudaSupport,
                          "HIP runtime/device not available")
        return self.init_context(0)

    def init_context(self, device: int = 0) -> Result:
        """Create the C-ABI context without probing for a device (device work
        is deferred to the first step, as in the reference)."""
        lib = load_library()
        self.clear()
        ctx = C.c_void_p()
        st = lib.silk_hip_create(device, C.byref(ctx))
        if st != 0:
            return Result(False, ErrorCode(st - 1), "context creation failed")
        self._ctx = ctx
        return Result()

    def _wrap(self, st: int) -> Result:
        if st == 0:
            return Result()
        detail = load_library().silk_hip_last_error(self._ctx) or b""
        return Result(False, ErrorCode(st - 1), detail.decode())

    def _need(self):
        if self._ctx is None:
            return Result(False, ErrorCode.NotInitialized)
        return None

    def set_global_config(self, config: GlobalConfig) -> Result:
        if (r := self._need()) is not None:
            return r
        g = config._c()
        return self._wrap(load_library().silk_hip_set_global(self._ctx, C.byref(g)))

    def clear(self) -> None:
        if getattr(self, "_ctx", None) is not None and _lib is not None:
            _lib.silk_hip_destroy(self._ctx)
        self._ctx = None

    # -- solver
    def solver_step(self) -> Result:
        if (r := self._need()) is not None:
            return r
        self.last_stats = SilkHipStepStats()
        return self._wrap(load_library().silk_hip_step(self._ctx,
                                                       C.byref(self.last_stats)))

    def solver_reset(self) -> Result:
        if (r := self._need()) is not None:
            return r
        return self._wrap(load_library().silk_hip_reset(self._ctx))

    # -- cloth
    def add_cloth(self, cloth_config: ClothConfig, collision_config: CollisionConfig,
                  verts, faces, pin_index=()) -> tuple[Result, int]:
        if (r := self._need()) is not None:
            return r, 0
        v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1)
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
        p = np.ascontiguousarray(pin_index, dtype=np.int32).reshape(-1)
        if v.size % 3 or f.size % 3:
            return Result(False, ErrorCode.InvalidMesh), 0
        h = C.c_uint32(0)
        cp, cc = cloth_config._c(), collision_config._c()
        st = load_library().silk_hip_add_cloth(
            self._ctx, C.byref(cp), C.byref(cc), _fptr(v), v.size // 3, _iptr(f),
            f.size // 3, _iptr(p), p.size, C.byref(h))
        return self._wrap(st), h.value

    def remove_cloth(self, handle: int) -> Result:
        if (r := self._need()) is not None:
            return r
        return self._wrap(load_library().silk_hip_remove(self._ctx, handle, 1))

    def get_cloth_position(self, handle: int, position: np.ndarray | None = None,
                           n: int | None = None):
        """Returns (Result, positions).  `position` (float32, size 3*nv) is
        filled in place when given; otherwise `n` floats are allocated."""
        if (r := self._need()) is not None:
            return r, None
        if position is None:
            if n is None:
                raise ValueError("pass a buffer or its size")
            position = np.empty(n, dtype=np.float32)
        st = load_library().silk_hip_get_positions(self._ctx, handle,
                                                   _fptr(position), position.size)
        return self._wrap(st), position

    def set_cloth_config(self, handle: int, config: ClothConfig) -> Result:
        if (r := self._need()) is not None:
            return r
        cp = config._c()
        return self._wrap(load_library().silk_hip_set_cloth_params(
            self._ctx, handle, C.byref(cp)))

    def set_cloth_collision_config(self, handle: int, config: CollisionConfig) -> Result:
        if (r := self._need()) is not None:
            return r
        cc = config._c()
        return self._wrap(load_library().silk_hip_set_collision_params(
            self._ctx, handle, 1, C.byref(cc)))

    def set_cloth_pin_index(self, handle: int, pin_index) -> Result:
        if (r := self._need()) is not None:
            return r
        p = np.ascontiguousarray(pin_index, dtype=np.int32).reshape(-1)
        return self._wrap(load_library().silk_hip_set_pin_index(
            self._ctx, handle, _iptr(p), p.size))

    def set_cloth_pin_position(self, handle: int, position) -> Result:
        if (r := self._need()) is not None:
            return r
        p = np.ascontiguousarray(position, dtype=np.float32).reshape(-1)
        return self._wrap(load_library().silk_hip_set_pin_positions(
            self._ctx, handle, _fptr(p), p.size))

    # -- obstacle
    def add_obstacle(self, collision_config: CollisionConfig, verts,
                     faces) -> tuple[Result, int]:
        if (r := self._need()) is not None:
            return r, 0
        v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1)
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
        if v.size % 3 or f.size % 3:
            return Result(False, ErrorCode.InvalidMesh), 0
        h = C.c_uint32(0)
        cc = collision_config._c()
        st = load_library().silk_hip_add_obstacle(
            self._ctx, C.byref(cc), _fptr(v), v.size // 3, _iptr(f), f.size // 3,
            C.byref(h))
        return self._wrap(st), h.value

    def remove_obstacle(self, handle: int) -> Result:
        if (r := self._need()) is not None:
            return r
        return self._wrap(load_library().silk_hip_remove(self._ctx, handle, 0))

    def set_obstacle_collision_config(self, handle: int,
                                      config: CollisionConfig) -> Result:
        if (r := self._need()) is not None:
            return r
        cc = config._c()
        return self._wrap(load_library().silk_hip_set_collision_params(
            self._ctx, handle, 0, C.byref(cc)))

    def set_obstacle_position(self, handle: int, position) -> Result:
        if (r := self._need()) is not None:
            return r
        p = np.ascontiguousarray(position, dtype=np.float32).reshape(-1)
        return self._wrap(load_library().silk_hip_set_obstacle_positions(
            self._ctx, handle, _fptr(p), p.size))

    # -- introspection (no reference counterpart)
    def get_partition(self, handle: int, nv: int):
        perm = np.empty(nv, dtype=np.int32)
        npart = C.c_int32(0)
        lib = load_library()
        st = lib.silk_hip_get_partition(self._ctx, handle, _iptr(perm), nv, None,
                                        0, C.byref(npart))
        if st != 0:
            return self._wrap(st), None, None
        offsets = np.empty(npart.value + 1, dtype=np.int32)
        st = lib.silk_hip_get_partition(self._ctx, handle, None, 0, _iptr(offsets),
                                        offsets.size, None)
        return self._wrap(st), perm, offsets

    def set_timing(self, on) -> Result:
        """False/True: off/every class; an int other than 0 and 1 is a mask of
        timing classes (bit c = class c)."""
        return self._wrap(load_library().silk_hip_set_timing(self._ctx, int(on)))

    SCENE_INFO_KEYS = ("vertices", "triangles", "offdiag_entries", "mas_banks",
                       "slots", "pcg_workgroups", "coarse_banks",
                       "collider_vertices", "pcg_launches", "systems",
                       "pcg_lds_bytes", "pcg_dense")

    def scene_info(self) -> dict:
        k = len(self.SCENE_INFO_KEYS)
        info = (C.c_int64 * k)()
        st = load_library().silk_hip_get_scene_info(self._ctx, info, k)
        if st != 0:
            raise RuntimeError(self._wrap(st).to_string())
        return dict(zip(self.SCENE_INFO_KEYS, [int(v) for v in info]))

    def collider_vertex_count(self) -> int:
        n = C.c_int64(0)
        st = load_library().silk_hip_collider_vertex_count(self._ctx, C.byref(n))
        if st != 0:
            raise RuntimeError(self._wrap(st).to_string())
        return n.value

    def debug_candidates(self, pos0, pos1):
        """Broadphase of the current scene on explicit collider positions:
        (PT tuples (p, t0, t1, t2), EE tuples (a0, a1, b0, b1))."""
        a = np.ascontiguousarray(pos0, dtype=np.float32).reshape(-1)
        b = np.ascontiguousarray(pos1, dtype=np.float32).reshape(-1)
        lib = load_library()
        npt, nee = C.c_int64(0), C.c_int64(0)
        st = lib.silk_hip_debug_candidates(self._ctx, _fptr(a), _fptr(b), a.size,
                                           None, 0, C.byref(npt), None, 0,
                                           C.byref(nee))
        if st != 0:
            raise RuntimeError(self._wrap(st).to_string())
        pt = np.zeros((max(1, npt.value), 4), dtype=np.int32)
        ee = np.zeros((max(1, nee.value), 4), dtype=np.int32)
        st = lib.silk_hip_debug_candidates(self._ctx, _fptr(a), _fptr(b), a.size,
                                           _iptr(pt), npt.value, C.byref(npt),
                                           _iptr(ee), nee.value, C.byref(nee))
        if st != 0:
            raise RuntimeError(self._wrap(st).to_string())
        return pt[:npt.value], ee[:nee.value]

    def debug_copy(self, nbytes: int, reps: int = 1) -> float:
        """Device copy of nbytes, reps times, with the library's copy kernel;
        returns the elapsed device time (ms)."""
        ms = C.c_double(0.0)
        st = load_library().silk_hip_debug_copy(self._ctx, nbytes, reps, C.byref(ms))
        if st != 0:
            raise RuntimeError(self._wrap(st).to_string())
        return ms.value

    def debug_dcd(self, pos0, pos1, time, dhat):
        """DCD of the current scene on explicit collider positions:
        (records (n, 5): type, four global vertex ids in record order;
        distances (n,))."""
        a = np.ascontiguousarray(pos0, dtype=np.float32).reshape(-1)
        b = np.ascontiguousarray(pos1, dtype=np.float32).reshape(-1)
        lib = load_library()
        n = C.c_int64(0)
        st = lib.silk_hip_debug_dcd(self._ctx, _fptr(a), _fptr(b), a.size, time, dhat,
                                    None, None, 0, C.byref(n))
        if st != 0:
            raise RuntimeError(self._wrap(st).to_string())
        recs = np.zeros((max(1, n.value), 6), dtype=np.int32)
        dist = np.zeros(max(1, n.value), dtype=np.float32)
        st = lib.silk_hip_debug_dcd(self._ctx, _fptr(a), _fptr(b), a.size, time, dhat,
                                    _iptr(recs), _fptr(dist), n.value, C.byref(n))
        if st != 0:
            raise RuntimeError(self._wrap(st).to_string())
        return recs[:n.value, :5], dist[:n.value]

    def decisions(self):
        """Discrete decisions of the last step: (TOI of every outer
        iteration, float32; appended contacts (n, 6): DCD call, type, four
        global vertex ids in record order; ADMM iterations of every outer
        iteration; per contact colour << 4 | mask of its vertices in the
        contact solve)."""
        lib = load_library()
        nt, nc = C.c_int32(0), C.c_int64(0)
        st = lib.silk_hip_debug_decisions(self._ctx, None, None, 0, C.byref(nt), None, None,
                                          0, C.byref(nc))
        if st != 0:
            raise RuntimeError(self._wrap(st).to_string())
        tois = np.zeros(max(1, nt.value), dtype=np.float32)
        admm = np.zeros(max(1, nt.value), dtype=np.int32)
        recs = np.zeros((max(1, nc.value), 6), dtype=np.int32)
        info = np.zeros(max(1, nc.value), dtype=np.int32)
        st = lib.silk_hip_debug_decisions(self._ctx, _fptr(tois), _iptr(admm), nt.value,
                                          C.byref(nt), _iptr(recs), _iptr(info), nc.value,
                                          C.byref(nc))
        if st != 0:
            raise RuntimeError(self._wrap(st).to_string())
        return tois[:nt.value], recs[:nc.value], admm[:nt.value], info[:nc.value]

    def debug_ccd(self, is_vf, p0, p1, err, ms, tolerance, max_iter=4096):
        """The step's FP64 CCD kernel on explicit queries: (hit, toi)."""
        v = np.ascontiguousarray(is_vf, dtype=np.int32).reshape(-1)
        a = np.ascontiguousarray(p0, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(p1, dtype=np.float64).reshape(-1)
        e = np.ascontiguousarray(err, dtype=np.float64).reshape(-1)
        n = v.size
        assert a.size == 12 * n and b.size == 12 * n and e.size == 3 * n
        hit = np.zeros(n, dtype=np.int32)
        toi = np.zeros(n, dtype=np.float64)
        dp = C.POINTER(C.c_double)
        st = load_library().silk_hip_debug_ccd(
            self._ctx, n, _iptr(v), a.ctypes.data_as(dp), b.ctypes.data_as(dp),
            e.ctypes.data_as(dp), ms, tolerance, max_iter, _iptr(hit),
            toi.ctypes.data_as(dp))
        if st != 0:
            raise RuntimeError(self._wrap(st).to_string())
        return hit, toi

    def debug_ccd_min(self, is_vf, p0, p1, err, ms, tolerance, max_iter=4096):
        """The step's earliest-TOI CCD kernel on explicit queries."""
        v = np.ascontiguousarray(is_vf, dtype=np.int32).reshape(-1)
        a = np.ascontiguousarray(p0, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(p1, dtype=np.float64).reshape(-1)
        e = np.ascontiguousarray(err, dtype=np.float64).reshape(-1)
        n = v.size
        assert a.size == 12 * n and b.size == 12 * n and e.size == 3 * n
        out = C.c_double(0.0)
        dp = C.POINTER(C.c_double)
        st = load_library().silk_hip_debug_ccd_min(
            self._ctx, n, _iptr(v), a.ctypes.data_as(dp), b.ctypes.data_as(dp),
            e.ctypes.data_as(dp), ms, tolerance, max_iter, C.byref(out))
        if st != 0:
            raise RuntimeError(self._wrap(st).to_string())
        return out.value

    def debug_coulomb(self, u, normal, mu):
        """The contact solve's Coulomb projection on explicit cases (n x 3
        velocities and normals, n friction coefficients); returns n x 3."""
        a = np.ascontiguousarray(u, dtype=np.float32).reshape(-1)
        b = np.ascontiguousarray(normal, dtype=np.float32).reshape(-1)
        m = np.ascontiguousarray(mu, dtype=np.float32).reshape(-1)
        n = m.size
        assert a.size == 3 * n and b.size == 3 * n
        out = np.zeros(3 * n, dtype=np.float32)
        st = load_library().silk_hip_debug_coulomb(
            self._ctx, n, a.ctypes.data, b.ctypes.data, m.ctypes.data, out.ctypes.data)
        if st != 0:
            raise RuntimeError(self._wrap(st).to_string())
        return out.reshape(n, 3)

    # -- launch autotuning (library extension, see include/silk_hip.h)
    def autotune(self) -> tuple[Result, dict]:
        """Time the broadphase split depths and the persistent PCG variants
        on the current scene (after one solver_step) and keep the fastest.
        Returns the result and the report: chosen depth and PCG variant,
        microseconds per broadphase launch by depth and per PCG solve by
        variant (-1: not timed)."""
        rep = np.full(AUTOTUNE_REPORT, -1, dtype=np.int32)
        r = self._wrap(load_library().silk_hip_autotune(
            self._ctx, rep.ctypes.data_as(C.POINTER(C.c_int32)), rep.size))
        return r, {"split_depth": int(rep[0]), "pcg_variant": int(rep[1]),
                   "broadphase_us": {d: int(rep[d]) for d in range(2, 8)},
                   "pcg_us": {0: int(rep[8]), 1: int(rep[9])}}

    # -- learned warm start (library extension, see include/silk_hip.h)
    def set_warm_start_net(self, params) -> Result:
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        return self._wrap(load_library().silk_hip_set_warm_start_net(
            self._ctx, _fptr(p), p.size))

    def get_warm_start_net(self) -> tuple[Result, np.ndarray]:
        p = np.zeros(WARM_NET_PARAMS, dtype=np.float32)
        r = self._wrap(load_library().silk_hip_get_warm_start_net(
            self._ctx, _fptr(p), p.size))
        return r, p

    def warm_start_record(self, capacity: int, rate: float) -> Result:
        return self._wrap(load_library().silk_hip_warm_start_record(
            self._ctx, int(capacity), float(rate)))

    def warm_start_samples(self) -> int:
        n = C.c_int64(0)
        load_library().silk_hip_warm_start_samples(self._ctx, C.byref(n))
        return n.value

    def warm_start_train(self, steps: int, batch: int, lr: float,
                         seed: int = 0) -> tuple[Result, np.ndarray]:
        loss = np.zeros(max(1, steps), dtype=np.float32)
        r = self._wrap(load_library().silk_hip_warm_start_train(
            self._ctx, int(steps), int(batch), float(lr), int(seed), _fptr(loss)))
        return r, loss[:steps]

    def debug_pcg(self, d, b, rel_tol: float):
        """One PCG solve of K + diag(d) on b from zero (solver order):
        returns (x, iterations)."""
        dd = np.ascontiguousarray(d, dtype=np.float32).reshape(-1)
        bb = np.ascontiguousarray(b, dtype=np.float32).reshape(-1)
        x = np.zeros_like(bb)
        it = C.c_int32(0)
        r = self._wrap(load_library().silk_hip_debug_pcg(
            self._ctx, _fptr(dd), dd.size, _fptr(bb), bb.size, float(rel_tol),
            _fptr(x), C.byref(it)))
        assert r, r.to_string()
        return x, it.value

    def debug_warm_net(self, feat, target):
        f = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, 16)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(-1)
        y = np.zeros(t.size, dtype=np.float32)
        g = np.zeros(WARM_NET_PARAMS, dtype=np.float32)
        loss = C.c_double(0.0)
        r = self._wrap(load_library().silk_hip_debug_warm_net(
            self._ctx, t.size, _fptr(f), _fptr(t), _fptr(y), _fptr(g),
            C.byref(loss)))
        assert r, r.to_string()
        return y, g, loss.value

    def timing(self, cls: int, reset: bool = False) -> tuple[float, int]:
        ms = C.c_double(0)
        n = C.c_int64(0)
        load_library().silk_hip_get_timing(self._ctx, cls, C.byref(ms), C.byref(n),
                                           int(reset))
        return ms.value, n.value


__all__ = ["World", "GlobalConfig", "ClothConfig", "CollisionConfig", "Result",
           "ErrorCode", "Backend", "load_library", "lib_path"]
