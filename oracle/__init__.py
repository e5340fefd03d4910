"""ctypes wrapper of the FP64 CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the reported CPU
baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "liboracle.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "-j8"], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        d, i, f = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_float)
        vp = C.c_void_p
        sig = {
            "oracle_create": (vp, []),
            "oracle_destroy": (None, [vp]),
            "oracle_set_global": (None, [vp, d]),
            "oracle_add_cloth": (C.c_int, [vp, d, d, f, C.c_int, i, C.c_int, i,
                                           C.c_int, i, i, C.c_int]),
            "oracle_finalize": (None, [vp]),
            "oracle_add_obstacle": (C.c_int, [vp, d, f, C.c_int, i, C.c_int]),
            "oracle_set_obstacle_position": (None, [vp, C.c_int, f]),
            "oracle_step": (C.c_int, [vp, d]),
            "oracle_get_positions": (None, [vp, C.c_int, d]),
            "oracle_set_state": (None, [vp, C.c_int, d, d]),
            "oracle_mass_area": (None, [d, C.c_int, i, C.c_int, d, d]),
            "oracle_jacobians": (None, [d, C.c_int, i, C.c_int, d]),
            "oracle_operator_dense": (None, [d, C.c_int, i, C.c_int, C.c_double,
                                             C.c_double, d, d]),
            "oracle_project_stretch": (None, [d, C.c_double, C.c_double, d]),
            "oracle_local_step": (None, [d, C.c_int, i, C.c_int, C.c_double, d, d,
                                         d, d, d, d, d]),
            "oracle_mas_apply": (None, [d, C.c_int, i, C.c_int, C.c_double,
                                        C.c_double, i, C.c_int, d, d, d, i]),
            "oracle_pcg": (None, [d, C.c_int, i, C.c_int, C.c_double, C.c_double, i,
                                  C.c_int, d, d, d, C.c_double, C.c_double,
                                  C.c_int, d, C.c_int]),
            "oracle_overlap_pairs": (C.c_int, [f, C.c_int, f, C.c_int, C.c_int, i,
                                               C.c_int]),
            "oracle_ccd": (C.c_int, [C.c_int, d, d, C.c_double, C.c_double,
                                     C.c_int, d, C.c_double, C.c_int, C.c_int, d]),
            "oracle_ccd_f32": (C.c_int, [C.c_int, f, f, C.c_float, C.c_float,
                                         C.c_int, f, C.c_float, f]),
            "oracle_ccd_filter": (None, [d, C.c_int, d]),
            "oracle_coulomb": (None, [d, d, C.c_double, d]),
            "oracle_ee_uv": (C.c_int, [d, C.c_double, d]),
            "oracle_candidates": (C.c_int, [vp, d, d, i, C.c_int, i, i, C.c_int,
                                            i]),
            "oracle_dcd_contact": (C.c_int, [C.c_int, d, d, i, d, C.c_double, d]),
            "oracle_set_replay": (None, [vp, d, i, C.c_int, i, i, C.c_int]),
            "oracle_set_obstacle_frame": (None, [vp, C.c_int, f, f]),
            "oracle_decisions": (C.c_int, [vp, d, i, C.c_int, i, i, i, C.c_int, i]),
            "oracle_dcd": (C.c_int, [vp, d, d, C.c_double, C.c_double, i, d, C.c_int,
                                     i]),
            "oracle_sym3_min_eig": (C.c_double, [d]),
            "oracle_contact_sweeps": (None, [C.c_int, i, d, d, d, d, d, C.c_int, d,
                                             d, d, d, C.c_int, d, i]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int))


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


class OracleWorld:
    """FP64 restatement of the GPU-path step.  add_cloth needs the solver
    partition (perm, offsets) used by the library under test."""

    def __init__(self, g: dict):
        self._w = lib().oracle_create()
        p = np.array([g.get("ax", 0.0), g.get("ay", 0.0), g.get("az", -10.0),
                      g.get("max_outer", 100), g.get("max_inner", 100),
                      g.get("dt", 1 / 60), g.get("lin_abs", 1e-7),
                      g.get("lin_rel_min", 1e-6), g.get("lin_rel_max", 1e-3),
                      g.get("lin_rel_init", 1e-2), g.get("admm_abs", 1e-5),
                      g.get("admm_rel", 1e-3), g.get("pcg_max_iter", 1000),
                      g.get("sweeps", 10), g.get("ccd_max_iter", 4096),
                      g.get("ccd_tol", 1e-6), g.get("ms_scale", 0.2),
                      g.get("dcd_scale", 1.0), g.get("pcg_criterion", 0),
                      g.get("ccd_zero", 0)],
                     dtype=np.float64)
        self._keep = [p]
        lib().oracle_set_global(self._w, _d(p)[1])
        self.sizes = []

    def __del__(self):
        if getattr(self, "_w", None):
            lib().oracle_destroy(self._w)
            self._w = None

    def add_cloth(self, V, F, pins, perm, offsets, k=100.0, kb=1e-4, density=0.1,
                  damping=0.01, coll=(1, 1, 0, 0.3, 0.3)):
        cl, clp = _d([k, kb, density, damping])
        co, cop = _d(coll)
        v, vp = _f(V)
        f, fp = _i(F)
        pn, pnp = _i(pins if len(pins) else [0])
        pm, pmp = _i(perm)
        of, ofp = _i(offsets)
        idx = lib().oracle_add_cloth(self._w, clp, cop, vp, v.size // 3, fp,
                                     f.size // 3, pnp, len(pins), pmp, ofp,
                                     of.size - 1)
        self.sizes.append(v.size)
        return idx

    def add_obstacle(self, V, F, coll=(1, 1, 0, 0.3, 0.3)):
        co, cop = _d(coll)
        v, vp = _f(V)
        f, fp = _i(F)
        return lib().oracle_add_obstacle(self._w, cop, vp, v.size // 3, fp,
                                         f.size // 3)

    def set_state(self, idx: int, x, v):
        """Overwrite cloth `idx`'s position and velocity (caller order)."""
        a, ap = _d(np.asarray(x).reshape(-1))
        b, bp = _d(np.asarray(v).reshape(-1))
        lib().oracle_set_state(self._w, idx, ap, bp)

    def set_obstacle_position(self, idx: int, V):
        v, vp = _f(V)
        lib().oracle_set_obstacle_position(self._w, idx, vp)

    def set_obstacle_frame(self, idx: int, prev, curr):
        """Obstacle idx moves from prev to curr over the next step."""
        a, ap = _f(prev)
        b, bp = _f(curr)
        lib().oracle_set_obstacle_frame(self._w, idx, ap, bp)

    def set_replay(self, tois, contacts, admm=None, info=None):
        """Replay another solver's decisions in the next step: tois[k] is the
        CCD result of outer iteration k (frame-time units) and admm[k] (when
        given) the number of ADMM iterations it ran; contacts (n, 6) holds
        (DCD call, type, four global vertex ids in record order) of every
        contact to append and info[i] (when given) contact i's colour << 4 |
        mask of its vertices in the contact solve."""
        t, tp = _d(np.asarray(tois, dtype=np.float64).reshape(-1))
        c, cp = _i(np.asarray(contacts, dtype=np.int32).reshape(-1, 6))
        ap = fp = None
        if admm is not None:
            a, ap = _i(np.asarray(admm, dtype=np.int32).reshape(-1))
            assert a.size == t.size
        if info is not None:
            f, fp = _i(np.asarray(info, dtype=np.int32).reshape(-1))
            assert f.size == c.shape[0]
        lib().oracle_set_replay(self._w, tp, ap, t.size, cp, fp, c.shape[0])

    def decisions(self):
        """Decisions of the last step: (TOIs, contacts (n, 6), replayed
        contacts whose record could not be formed, ADMM iterations of every
        outer iteration, per contact colour << 4 | vertex mask)."""
        nt, nc = C.c_int(0), C.c_int(0)
        lib().oracle_decisions(self._w, None, None, 0, C.byref(nt), None, None, 0,
                               C.byref(nc))
        t = np.zeros(max(1, nt.value))
        a = np.zeros(max(1, nt.value), dtype=np.int32)
        c = np.zeros((max(1, nc.value), 6), dtype=np.int32)
        f = np.zeros(max(1, nc.value), dtype=np.int32)
        ip = C.POINTER(C.c_int)
        miss = lib().oracle_decisions(self._w, t.ctypes.data_as(C.POINTER(C.c_double)),
                                      a.ctypes.data_as(ip), nt.value, C.byref(nt),
                                      c.ctypes.data_as(ip), f.ctypes.data_as(ip), nc.value,
                                      C.byref(nc))
        return t[:nt.value], c[:nc.value], miss, a[:nt.value], f[:nc.value]

    def dcd(self, pos0, pos1, time, dhat):
        """DCD records on explicit collider positions (global numbering):
        (records (n, 5): type, four ids in record order; distances (n,))."""
        a, ap = _d(pos0)
        b, bp = _d(pos1)
        n = C.c_int(0)
        cap = 1 << 14
        while True:
            recs = np.zeros((cap, 6), dtype=np.int32)
            dist = np.zeros(cap)
            rc = lib().oracle_dcd(self._w, ap, bp, time, dhat,
                                  recs.ctypes.data_as(C.POINTER(C.c_int)),
                                  dist.ctypes.data_as(C.POINTER(C.c_double)), cap,
                                  C.byref(n))
            if rc == 0:
                return recs[:n.value, :5], dist[:n.value]
            cap = n.value

    def candidates(self, pos0, pos1):
        """Broadphase candidates for explicit collider positions (global
        numbering): (PT tuples (p, t0, t1, t2), EE tuples (a0, a1, b0, b1))."""
        a, ap = _d(pos0)
        b, bp = _d(pos1)
        npt, nee = C.c_int(0), C.c_int(0)
        cap_pt, cap_ee = 1 << 16, 1 << 16
        while True:
            pt = np.zeros(4 * cap_pt, dtype=np.int32)
            ee = np.zeros(4 * cap_ee, dtype=np.int32)
            rc = lib().oracle_candidates(
                self._w, ap, bp, pt.ctypes.data_as(C.POINTER(C.c_int)), cap_pt,
                C.byref(npt), ee.ctypes.data_as(C.POINTER(C.c_int)), cap_ee,
                C.byref(nee))
            if rc == 0:
                return (pt[:4 * npt.value].reshape(-1, 4),
                        ee[:4 * nee.value].reshape(-1, 4))
            cap_pt, cap_ee = max(cap_pt, npt.value), max(cap_ee, nee.value)

    def finalize(self):
        lib().oracle_finalize(self._w)

    def step(self) -> tuple[int, dict]:
        st = np.zeros(8)
        rc = lib().oracle_step(self._w, st.ctypes.data_as(C.POINTER(C.c_double)))
        keys = ["outer", "admm", "pcg", "contacts", "colors", "ccd_candidates",
                "min_toi", "ccd_zero"]
        return rc, dict(zip(keys, st.tolist()))

    def positions(self, idx: int) -> np.ndarray:
        out = np.zeros(self.sizes[idx], dtype=np.float64)
        lib().oracle_get_positions(self._w, idx,
                                   out.ctypes.data_as(C.POINTER(C.c_double)))
        return out


# ------------------------------------------------------------ unit wrappers

def mass_area(V, F):
    v, vp = _d(V)
    f, fp = _i(F)
    m = np.zeros(v.size // 3)
    a = np.zeros(f.size // 3)
    lib().oracle_mass_area(vp, v.size // 3, fp, f.size // 3, _d(m)[1] if False else
                           m.ctypes.data_as(C.POINTER(C.c_double)),
                           a.ctypes.data_as(C.POINTER(C.c_double)))
    return m, a


def jacobians(V, F):
    v, vp = _d(V)
    f, fp = _i(F)
    J = np.zeros((f.size // 3, 6, 9))
    lib().oracle_jacobians(vp, v.size // 3, fp, f.size // 3,
                           J.ctypes.data_as(C.POINTER(C.c_double)))
    return J


def operator_dense(V, F, k, kb):
    v, vp = _d(V)
    f, fp = _i(F)
    n3 = v.size
    A = np.zeros((n3, n3))
    b = np.zeros(n3)
    lib().oracle_operator_dense(vp, v.size // 3, fp, f.size // 3, k, kb,
                                A.ctypes.data_as(C.POINTER(C.c_double)),
                                b.ctypes.data_as(C.POINTER(C.c_double)))
    return A, b


def project_stretch(y, lo=0.9, hi=1.1):
    y, yp = _d(y)
    out = np.zeros(6)
    lib().oracle_project_stretch(yp, lo, hi, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def local_step(V, F, k, x, u, z):
    v, vp = _d(V)
    f, fp = _i(F)
    xx, xp = _d(x)
    uu = np.array(u, dtype=np.float64)
    zz = np.array(z, dtype=np.float64)
    nv = v.size // 3
    sums = np.zeros(3)
    dual, dcur, dprev = np.zeros(3 * nv), np.zeros(3 * nv), np.zeros(3 * nv)
    P = C.POINTER(C.c_double)
    lib().oracle_local_step(vp, nv, fp, f.size // 3, k, xp, uu.ctypes.data_as(P),
                            zz.ctypes.data_as(P), sums.ctypes.data_as(P),
                            dual.ctypes.data_as(P), dcur.ctypes.data_as(P),
                            dprev.ctypes.data_as(P))
    return uu, zz, sums, dual, dcur, dprev


def mas_apply(V, F, k, kb, offsets, diag, r):
    v, vp = _d(V)
    f, fp = _i(F)
    o, op = _i(offsets)
    dg, dgp = _d(diag)
    rr, rp = _d(r)
    z = np.zeros(v.size)
    lv = C.c_int(0)
    lib().oracle_mas_apply(vp, v.size // 3, fp, f.size // 3, k, kb, op, o.size - 1,
                           dgp, rp, z.ctypes.data_as(C.POINTER(C.c_double)),
                           C.byref(lv))
    return z, lv.value


def pcg(V, F, k, kb, offsets, diag, b, x0, rel_tol, abs_tol, max_iter=1000,
        criterion=0):
    v, vp = _d(V)
    f, fp = _i(F)
    o, op = _i(offsets)
    dg, dgp = _d(diag)
    bb, bp = _d(b)
    x = np.array(x0, dtype=np.float64)
    out = np.zeros(4)
    lib().oracle_pcg(vp, v.size // 3, fp, f.size // 3, k, kb, op, o.size - 1, dgp,
                     bp, x.ctypes.data_as(C.POINTER(C.c_double)), rel_tol, abs_tol,
                     max_iter, out.ctypes.data_as(C.POINTER(C.c_double)),
                     int(criterion))
    return x, {"iters": int(out[0]), "status": int(out[1]), "rr0": out[2],
               "rr": out[3]}


def overlap_pairs(A, B=None):
    a, ap = _f(np.asarray(A).reshape(-1, 6))
    self = B is None
    b, bp = _f(np.asarray(A if self else B).reshape(-1, 6))
    n = lib().oracle_overlap_pairs(ap, a.size // 6, bp, b.size // 6, int(self),
                                   None, 0)
    out = np.zeros((max(n, 1), 2), dtype=np.int32)
    lib().oracle_overlap_pairs(ap, a.size // 6, bp, b.size // 6, int(self),
                               out.ctypes.data_as(C.POINTER(C.c_int)), n)
    return out[:n]


def ccd(vf, p0, p1, ms, tolerance, err, max_iter=4096, max_time=1.0,
        stack_cap=1024, zero_mode=0, with_zero=False):
    """(hit, toi) of one FP64 CCD query; with_zero adds whether the last
    search still had a root at t = 0 (reported as no hit)."""
    a, ap = _d(p0)
    b, bp = _d(p1)
    e, ep = _d(err)
    t = C.c_double(0)
    rc = lib().oracle_ccd(int(vf), ap, bp, ms, tolerance, max_iter, ep, max_time,
                          stack_cap, int(zero_mode), C.byref(t))
    if with_zero:
        return rc == 1, t.value, rc == 2
    return rc == 1, t.value


def ccd_f32(vf, p0, p1, ms, tolerance, err, max_iter=4096, max_time=1.0):
    a, ap = _f(p0)
    b, bp = _f(p1)
    e, ep = _f(err)
    t = C.c_float(0)
    hit = lib().oracle_ccd_f32(int(vf), ap, bp, ms, tolerance, max_iter, ep,
                               max_time, C.byref(t))
    return bool(hit), t.value


def ccd_filter(abs_max, vf):
    a, ap = _d(abs_max)
    e = np.zeros(3)
    lib().oracle_ccd_filter(ap, int(vf), e.ctypes.data_as(C.POINTER(C.c_double)))
    return e


def coulomb(u_star, normal, mu):
    a, ap = _d(u_star)
    n, np_ = _d(normal)
    out = np.zeros(3)
    lib().oracle_coulomb(ap, np_, mu, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def ee_uv(x, eps=1e-6):
    a, ap = _d(np.asarray(x).reshape(-1))
    out = np.zeros(2)
    ok = lib().oracle_ee_uv(ap, eps, out.ctypes.data_as(C.POINTER(C.c_double)))
    return bool(ok), out


def dcd_contact(pt: bool, x, inv_mass, is_static, cw, dhat):
    """DCD record and Daviet contact of one pair on explicit positions (4x3):
    None when the pair makes no contact, else a dict with dist, u, v, normal,
    coef, gamma, state (bool), affine and the contact's u(x)."""
    out = np.zeros(37)
    xs, xp = _d(np.asarray(x).reshape(-1))
    im, imp = _d(inv_mass)
    st, stp = _i(np.asarray(is_static, dtype=np.int32))
    c, cp = _d(cw)
    rc = lib().oracle_dcd_contact(int(pt), xp, imp, stp, cp, float(dhat),
                                  out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc == 0:
        return None
    return dict(dist=out[0], u=out[1], v=out[2], normal=out[3:6], coef=out[6:10],
                gamma=out[10:14], state=out[14:18] > 0.5, affine=out[18:21],
                u_at_x=out[21:24])


def sym3_min_eig(a) -> float:
    a, ap = _d(np.asarray(a).reshape(9))
    return float(lib().oracle_sym3_min_eig(ap))


def contact_sweeps(state, coef, gamma, normal, affine, mu, x, cw, eta=None,
                   corr=None, sweeps=10):
    """One ADMM contact update on explicit contacts; returns (p, eta, corr,
    colours)."""
    state = np.ascontiguousarray(state, dtype=np.int32).reshape(-1, 4)
    nc = state.shape[0]
    nv = np.asarray(x).size // 3
    eta = np.zeros(3 * nv) if eta is None else np.array(eta, dtype=np.float64).reshape(-1)
    corr = np.zeros(3 * nc) if corr is None else np.array(corr, dtype=np.float64).reshape(-1)
    p = np.zeros(3 * nv)
    col = np.zeros(nc, dtype=np.int32)
    keep = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            for a in (coef, gamma, normal, affine, mu, x, cw)]
    ptr = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in keep]
    lib().oracle_contact_sweeps(
        nc, state.ctypes.data_as(C.POINTER(C.c_int)), ptr[0], ptr[1], ptr[2], ptr[3],
        ptr[4], nv, ptr[5], ptr[6], eta.ctypes.data_as(C.POINTER(C.c_double)),
        corr.ctypes.data_as(C.POINTER(C.c_double)), int(sweeps),
        p.ctypes.data_as(C.POINTER(C.c_double)), col.ctypes.data_as(C.POINTER(C.c_int)))
    return p, eta, corr.reshape(nc, 3), col


# ----------------------------------------------- reference CPU pipeline

PIPE_LIB = HERE / "build" / "libcpu_pipeline.so"
_pipe = None


def pipe_lib():
    """The reference CPU algorithm restated (oracle/cpu_pipeline.cpp), built
    -O3 -march=native: the timing baseline of bench.py."""
    global _pipe
    if _pipe is None:
        if not PIPE_LIB.exists():
            build()
        L = C.CDLL(str(PIPE_LIB))
        d, i, f = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_float)
        vp = C.c_void_p
        sig = {
            "cpu_pipeline_create": (vp, [C.c_int, d]),
            "cpu_pipeline_destroy": (None, [vp]),
            "cpu_pipeline_add_cloth": (C.c_int, [vp, f, C.c_int, i, C.c_int, i, C.c_int]),
            "cpu_pipeline_add_obstacle": (C.c_int, [vp, f, C.c_int, i, C.c_int]),
            "cpu_pipeline_step": (C.c_int, [vp, d]),
            "cpu_pipeline_state": (None, [vp, C.c_int, f, f, C.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _pipe = L
    return _pipe


class CpuPipeline:
    """The reference CPU solver step (pipeline.cpp:45-168) restated: PD with
    a project-written sparse Cholesky factor, per-face SVD projection,
    barrier reflection targets and FP32 CCD rollback, on `threads` host
    threads.  Default parameters are the reference's ClothConfig and
    GlobalConfig defaults with dt 1/60 and g = -9.81 z."""

    def __init__(self, threads: int = 1, dt: float = 1 / 60, g=(0.0, 0.0, -9.81),
                 k: float = 100.0, kb: float = 1e-4, density: float = 0.1,
                 damping: float = 0.01, max_outer: int = 100, max_inner: int = 10,
                 ms_scale: float = 0.2, dcd_scale: float = 1.0, ccd_tol: float = 1e-6,
                 ccd_max_iter: int = 4096):
        p = np.array([dt, *g, k, kb, density, damping, max_outer, max_inner, ms_scale,
                      dcd_scale, ccd_tol, ccd_max_iter], dtype=np.float64)
        self._w = pipe_lib().cpu_pipeline_create(threads, p.ctypes.data_as(C.POINTER(C.c_double)))
        self.sizes = []

    def __del__(self):
        if getattr(self, "_w", None):
            pipe_lib().cpu_pipeline_destroy(self._w)
            self._w = None

    def add_cloth(self, V, F, pins=()):
        v, vp = _f(V)
        f, fp = _i(F)
        p, pp = _i(np.asarray(pins, dtype=np.int32).reshape(-1))
        rc = pipe_lib().cpu_pipeline_add_cloth(self._w, vp, v.size // 3, fp, f.size // 3,
                                               pp, p.size)
        self.sizes.append(v.size)
        return rc

    def add_obstacle(self, V, F):
        v, vp = _f(V)
        f, fp = _i(F)
        return pipe_lib().cpu_pipeline_add_obstacle(self._w, vp, v.size // 3, fp, f.size // 3)

    def step(self) -> tuple[int, dict]:
        st = np.zeros(5)
        rc = pipe_lib().cpu_pipeline_step(self._w, st.ctypes.data_as(C.POINTER(C.c_double)))
        return rc, dict(zip(["outer", "inner", "collisions", "ccd_candidates", "min_toi"],
                            st.tolist()))

    def positions(self, idx: int = 0) -> np.ndarray:
        x = np.zeros(self.sizes[idx], dtype=np.float32)
        v = np.zeros_like(x)
        pipe_lib().cpu_pipeline_state(self._w, idx, _f(x)[1] if False else
                                      x.ctypes.data_as(C.POINTER(C.c_float)),
                                      v.ctypes.data_as(C.POINTER(C.c_float)), 0)
        return x

    def set_state(self, idx: int, x, v):
        a, ap = _f(np.asarray(x).reshape(-1))
        b, bp = _f(np.asarray(v).reshape(-1))
        pipe_lib().cpu_pipeline_state(self._w, idx, ap, bp, 1)
