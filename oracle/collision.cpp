// Oracle: exact box-overlap sets and the interval root-finder CCD.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// The CCD follows silk/src/backend/cpu/collision/interval_root_finder.cpp and
// ccd.cpp: boxes over (t, u, v) are evaluated with the affine L1 distance
// graph, the pending boxes are processed earliest t.lower first (ties: the
// most recently pushed), and the first box that is small enough (co-domain
// width <= tolerance or image inside the eps+ms cube) gives the TOI.  The
// single-precision instantiation restates the reference as it runs; the
// double one is what the HIP kernel must reproduce bit for bit.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "oracle.hpp"

namespace oracle {

std::vector<std::pair<int, int>> overlap_pairs(const std::vector<Box>& A,
                                               const std::vector<Box>& B,
                                               bool self) {
  // Sort-and-sweep on x with closed intervals; every candidate is confirmed by
  // the full 3-D closed overlap test, so the result equals brute force.
  std::vector<std::pair<int, int>> out;
  auto sorted = [](const std::vector<Box>& v) {
    std::vector<int> idx(v.size());
    for (size_t i = 0; i < v.size(); ++i) idx[i] = static_cast<int>(i);
    std::sort(idx.begin(), idx.end(), [&v](int a, int b) {
      return v[a].mn[0] != v[b].mn[0] ? v[a].mn[0] < v[b].mn[0] : a < b;
    });
    return idx;
  };
  auto prune = [](std::vector<int>& act, const std::vector<Box>& v, float m) {
    size_t w = 0;
    for (int i : act)
      if (v[i].mx[0] >= m) act[w++] = i;
    act.resize(w);
  };
  if (self) {
    std::vector<int> ia = sorted(A), act;
    for (int a : ia) {
      prune(act, A, A[a].mn[0]);
      for (int b : act)
        if (box_overlap(A[a], A[b])) out.emplace_back(std::min(a, b), std::max(a, b));
      act.push_back(a);
    }
  } else {
    std::vector<int> ia = sorted(A), ib = sorted(B), actA, actB;
    size_t i = 0, j = 0;
    while (i < ia.size() || j < ib.size()) {
      const bool take_a =
          j >= ib.size() || (i < ia.size() && A[ia[i]].mn[0] <= B[ib[j]].mn[0]);
      if (take_a) {
        const int a = ia[i++];
        prune(actB, B, A[a].mn[0]);
        for (int b : actB)
          if (box_overlap(A[a], B[b])) out.emplace_back(a, b);
        actA.push_back(a);
      } else {
        const int b = ib[j++];
        prune(actA, A, B[b].mn[0]);
        for (int a : actA)
          if (box_overlap(A[a], B[b])) out.emplace_back(a, b);
        actB.push_back(b);
      }
    }
  }
  std::sort(out.begin(), out.end());
  return out;
}

namespace {

template <typename T>
struct Ival {
  T lo, hi;
};

template <typename T>
struct IBox {
  Ival<T> d[3];  // t, u, v
};

template <typename T>
struct Coef {
  T c[6][3];
};

template <typename T>
T linf(const T a[3], const T b[3]) {
  T m = 0;
  for (int i = 0; i < 3; ++i) m = std::max(m, std::fabs(b[i] - a[i]));
  return m;
}

template <typename T>
T clamp_div(T a, T b) {
  if (b == 0) return std::numeric_limits<T>::infinity();
  return std::min(a / b, std::numeric_limits<T>::infinity());
}

// ccd.cpp:139-193 (vertex-face and edge-edge parameter tolerances).
template <typename T>
void param_tolerance(bool vf, const T p0[4][3], const T p1[4][3], T dist,
                     T tol[3]) {
  T q[2][4][3];  // [time][corner 000,001,011,010 order][xyz]
  for (int t = 0; t < 2; ++t) {
    const T(*p)[3] = t == 0 ? p0 : p1;
    for (int d = 0; d < 3; ++d) {
      if (vf) {
        q[t][0][d] = p[0][d] - p[1][d];                        // v - f0
        q[t][1][d] = p[0][d] - p[3][d];                        // v - f2
        q[t][2][d] = p[0][d] - (p[2][d] + p[3][d] - p[1][d]);  // v - (f1+f2-f0)
        q[t][3][d] = p[0][d] - p[2][d];                        // v - f1
      } else {
        q[t][0][d] = p[0][d] - p[2][d];  // a0 - b0
        q[t][1][d] = p[0][d] - p[3][d];  // a0 - b1
        q[t][2][d] = p[1][d] - p[3][d];  // a1 - b1
        q[t][3][d] = p[1][d] - p[2][d];  // a1 - b0
      }
    }
  }
  // p000 = q[0][0], p001 = q[0][1], p011 = q[0][2], p010 = q[0][3], p1xx = q[1][.]
  auto m4 = [&](const T* a1, const T* b1, const T* a2, const T* b2,
                const T* a3, const T* b3, const T* a4, const T* b4) {
    return std::max(std::max(linf(a1, b1), linf(a2, b2)),
                    std::max(linf(a3, b3), linf(a4, b4)));
  };
  const T* p000 = q[0][0]; const T* p001 = q[0][1];
  const T* p011 = q[0][2]; const T* p010 = q[0][3];
  const T* p100 = q[1][0]; const T* p101 = q[1][1];
  const T* p111 = q[1][2]; const T* p110 = q[1][3];
  const T dl = 3 * m4(p000, p100, p001, p101, p011, p111, p010, p110);
  const T e0 = 3 * m4(p000, p010, p100, p110, p101, p111, p001, p011);
  const T e1 = 3 * m4(p000, p001, p100, p101, p110, p111, p010, p011);
  tol[0] = clamp_div(dist, dl);
  tol[1] = clamp_div(dist, e0);
  tol[2] = clamp_div(dist, e1);
}

template <typename T>
Coef<T> coefficients(bool vf, const T p0[4][3], const T p1[4][3]) {
  Coef<T> k;
  for (int d = 0; d < 3; ++d) {
    if (vf) {
      k.c[0][d] = p0[0][d] - p0[1][d];
      k.c[1][d] = p1[0][d] - p1[1][d];
      k.c[2][d] = p0[1][d] - p0[2][d];
      k.c[3][d] = p1[1][d] - p1[2][d];
      k.c[4][d] = p0[1][d] - p0[3][d];
      k.c[5][d] = p1[1][d] - p1[3][d];
    } else {
      k.c[0][d] = p0[0][d] - p0[2][d];
      k.c[1][d] = p1[0][d] - p1[2][d];
      k.c[2][d] = p0[1][d] - p0[0][d];
      k.c[3][d] = p1[1][d] - p1[0][d];
      k.c[4][d] = p0[2][d] - p0[3][d];
      k.c[5][d] = p1[2][d] - p1[3][d];
    }
  }
  return k;
}

// eval_unit_bbox_1d (interval_root_finder.cpp:31-100).
template <typename T>
bool unit_eval(bool vf, const T p0[4][3], const T p1[4][3], T eps_ms, int d,
               bool* in_eps, T* width) {
  T mn, mx;
  if (vf) {
    const T D[6] = {p0[0][d] - p0[1][d], p0[0][d] - p0[2][d], p0[0][d] - p0[3][d],
                    p1[0][d] - p1[1][d], p1[0][d] - p1[2][d], p1[0][d] - p1[3][d]};
    mn = *std::min_element(D, D + 6);
    mx = *std::max_element(D, D + 6);
  } else {
    const T D[8] = {p0[0][d] - p0[2][d], p0[0][d] - p0[3][d], p0[1][d] - p0[2][d],
                    p0[1][d] - p0[3][d], p1[0][d] - p1[2][d], p1[0][d] - p1[3][d],
                    p1[1][d] - p1[2][d], p1[1][d] - p1[3][d]};
    mn = *std::min_element(D, D + 8);
    mx = *std::max_element(D, D + 8);
  }
  *width = mx - mn;
  *in_eps = false;
  if (mn > eps_ms || mx < -eps_ms) return false;
  if (mn >= -eps_ms && mx <= eps_ms) *in_eps = true;
  return true;
}

// eval_vf_bbox_1d / eval_ee_bbox_1d (interval_root_finder.cpp:102-276).
template <typename T>
bool box_eval(bool vf, const IBox<T>& b, const Coef<T>& k, T eps_ms, int d,
              bool* in_eps, T* width) {
  const T c0 = k.c[0][d], c1 = k.c[1][d], c2 = k.c[2][d], c3 = k.c[3][d],
          c4 = k.c[4][d], c5 = k.c[5][d];
  const T ulo = b.d[1].lo, uhi = b.d[1].hi, vlo = b.d[2].lo, vhi = b.d[2].hi;
  auto at_t = [&](T t, T* lo, T* hi) {
    const T q0 = (c1 - c0) * t + c0;
    const T q1 = (c3 - c2) * t + c2;
    const T q2 = (c5 - c4) * t + c4;
    if (!vf || uhi + vhi <= T(1)) {
      const T a = q1 * ulo, bb = q1 * uhi, c = q2 * vlo, dd = q2 * vhi;
      *lo = q0 + std::min(a, bb) + std::min(c, dd);
      *hi = q0 + std::max(a, bb) + std::max(c, dd);
      return;
    }
    *lo = std::numeric_limits<T>::infinity();
    *hi = -std::numeric_limits<T>::infinity();
    auto inc = [&](T u, T v) {
      const T val = q0 + q1 * u + q2 * v;
      *lo = std::min(*lo, val);
      *hi = std::max(*hi, val);
    };
    if (ulo + vlo <= T(1)) inc(ulo, vlo);
    if (ulo + vhi <= T(1)) inc(ulo, vhi);
    if (uhi + vlo <= T(1)) inc(uhi, vlo);
    if (uhi + vhi <= T(1)) inc(uhi, vhi);
    const T v_at_ulo = T(1) - ulo;
    if (v_at_ulo >= vlo && v_at_ulo <= vhi) inc(ulo, v_at_ulo);
    const T v_at_uhi = T(1) - uhi;
    if (v_at_uhi >= vlo && v_at_uhi <= vhi) inc(uhi, v_at_uhi);
    const T u_at_vlo = T(1) - vlo;
    if (u_at_vlo >= ulo && u_at_vlo <= uhi) inc(u_at_vlo, vlo);
    const T u_at_vhi = T(1) - vhi;
    if (u_at_vhi >= ulo && u_at_vhi <= uhi) inc(u_at_vhi, vhi);
  };
  T lo0, hi0, lo1, hi1;
  at_t(b.d[0].lo, &lo0, &hi0);
  at_t(b.d[0].hi, &lo1, &hi1);
  const T mn = std::min(lo0, lo1), mx = std::max(hi0, hi1);
  *in_eps = false;
  if (mn > eps_ms || mx < -eps_ms) return false;
  *width = mx - mn;
  if (mn >= -eps_ms && mx <= eps_ms) *in_eps = true;
  return true;
}

struct Pending {
  int n = 0;
};

// interval_root_finder_bucket_DFS (interval_root_finder.cpp:391-510) with a
// bounded pending list; on overflow the current box is returned (a
// conservative, earlier TOI).  Returns true and the box t.lower on a root.
template <typename T>
bool root_find(bool vf, const T p0[4][3], const T p1[4][3], const T tol[3],
               T co_tol, const T err[3], T ms, long max_iter, T max_time,
               int cap, T* t_lo) {
  const Coef<T> k = coefficients(vf, p0, p1);
  std::vector<IBox<T>> stack;
  stack.reserve(64);
  IBox<T> cur{{{T(0), max_time}, {T(0), T(1)}, {T(0), T(1)}}};
  bool has_cur = true;
  bool unit = (max_time == T(1));
  long iter = 0;
  while (has_cur || !stack.empty()) {
    if (!has_cur) {
      // smallest t.lower; ties -> latest pushed (highest index)
      size_t best = stack.size() - 1;
      for (size_t i = stack.size() - 1; i-- > 0;)
        if (stack[i].d[0].lo < stack[best].d[0].lo) best = i;
      cur = stack[best];
      stack.erase(stack.begin() + static_cast<long>(best));
    }
    has_cur = false;
    ++iter;
    bool in_eps_all = true;
    bool hit = true;
    T width[3];
    for (int d = 0; d < 3; ++d) {
      bool in_eps = false;
      const T em = err[d] + ms;
      const bool ok = unit ? unit_eval(vf, p0, p1, err[d] + ms, d, &in_eps, &width[d])
                           : box_eval(vf, cur, k, em, d, &in_eps, &width[d]);
      if (!ok) {
        hit = false;
        break;
      }
      in_eps_all = in_eps_all && in_eps;
    }
    unit = false;
    if (!hit) continue;
    const bool small = (width[0] <= co_tol && width[1] <= co_tol &&
                        width[2] <= co_tol) || in_eps_all;
    if (small) {
      *t_lo = cur.d[0].lo;
      return true;
    }
    if (max_iter > 0 && iter > max_iter) {
      *t_lo = cur.d[0].lo;
      return true;
    }
    // find_next_split: largest width/tol among widths above their tolerance.
    int split = 0;
    T best = -std::numeric_limits<T>::infinity();
    for (int d = 0; d < 3; ++d) {
      const T w = cur.d[d].hi - cur.d[d].lo;
      const T s = (w > tol[d]) ? w / tol[d] : -std::numeric_limits<T>::infinity();
      if (s > best) {
        best = s;
        split = d;
      }
    }
    const T mid = (cur.d[split].hi + cur.d[split].lo) / 2;
    if (mid < T(1e-12)) {  // SPLITING_LOWER_BOUND on the first half's upper
      *t_lo = cur.d[0].lo;
      return true;
    }
    IBox<T> second = cur, first = cur;
    second.d[split] = {mid, cur.d[split].hi};
    first.d[split] = {cur.d[split].lo, mid};
    bool push_second = true, push_first = true;
    if (vf && split != 0) {
      const T other = (split == 1) ? cur.d[2].lo : cur.d[1].lo;
      push_second = second.d[split].lo + other <= T(1);
      push_first = first.d[split].lo + other <= T(1);
    }
    const int need = static_cast<int>(push_second) + static_cast<int>(push_first);
    if (static_cast<int>(stack.size()) + need > cap) {
      *t_lo = cur.d[0].lo;
      return true;
    }
    if (push_second) stack.push_back(second);
    if (push_first) stack.push_back(first);
  }
  return false;
}

template <typename T>
bool ccd_generic(bool vf, const T p0[4][3], const T p1[4][3], T ms, T tolerance,
                 long max_iter, const T err[3], T max_time, bool no_zero_toi,
                 int zero_mode, int cap, T* toi, bool* zero) {
  T tol[3];
  param_tolerance<T>(vf, p0, p1, tolerance, tol);
  T t;
  if (zero) *zero = false;
  if (!root_find<T>(vf, p0, p1, tol, tolerance, err, ms, max_iter, max_time, cap,
                    &t))
    return false;
  if (!no_zero_toi || t > T(0)) {
    *toi = t;
    return true;
  }
  // ccd.cpp:150-180: one refinement at ms/2 (zero_mode 1: then one at zero
  // separation).  No root there: no contact before max_time; a root at t = 0
  // again: no collision for the segment (narrowphase.cu:68).
  const int passes = zero_mode == 1 ? 2 : 1;
  for (int pass = 0; pass < passes; ++pass) {
    T t2;
    const T sep = pass == 0 ? ms * T(0.5) : T(0);
    if (!root_find<T>(vf, p0, p1, tol, tolerance, err, sep, max_iter, max_time, cap,
                      &t2)) {
      *toi = max_time;
      return true;
    }
    if (t2 > T(0)) {
      *toi = t2;
      return true;
    }
  }
  if (zero) *zero = true;
  return false;
}

}  // namespace

bool ccd_query(bool vf, const double p0[4][3], const double p1[4][3],
               const CcdConfig& cfg, double* toi, bool* zero) {
  return ccd_generic<double>(vf, p0, p1, cfg.ms, cfg.tolerance, cfg.max_iter,
                             cfg.err, cfg.max_time, cfg.no_zero_toi, cfg.zero_mode,
                             cfg.stack_cap, toi, zero);
}

bool ccd_query_f32(bool vf, const float p0[4][3], const float p1[4][3],
                   float ms, float tolerance, long max_iter, const float err[3],
                   float max_time, float* toi) {
  return ccd_generic<float>(vf, p0, p1, ms, tolerance, max_iter, err, max_time,
                            true, 0, 1 << 30, toi, nullptr);
}

void ccd_filter(const double abs_max[3], bool vf, double err[3]) {
  // Single-precision filters of interval_root_finder.cpp:516-517, applied to
  // both precisions (conservative for the double evaluation).
  const double f = vf ? static_cast<double>(4.053116e-06f)
                      : static_cast<double>(3.814698e-06f);
  for (int d = 0; d < 3; ++d) err[d] = f * std::max(abs_max[d], 1.0);
}

}  // namespace oracle
