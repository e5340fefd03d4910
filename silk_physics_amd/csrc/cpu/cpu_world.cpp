#include "cpu/cpu_world.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <queue>
#include <stdexcept>
#include <thread>

#include <sched.h>

#include "cpu/sparse_cholesky.hpp"

namespace silk::amd::cpu {

namespace {

constexpr double kPinPenalty = 1e4;        // cloth_solver_context.cpp:17
constexpr double kBarrierStiffness = 1e4;  // collision/pipeline.hpp:29
constexpr float kDcdEps = 1e-6f;           // narrowphase.cpp:9
constexpr double kSplitLowerBound = 1e-12;  // interval root finder split floor

// ------------------------------------------------------------ worker pool

// Persistent workers; run() hands out chunks of [0, n) through an atomic
// counter and the calling thread works too.
class Pool {
 public:
  explicit Pool(int n) {
    for (int i = 1; i < n; ++i) th_.emplace_back([this, i] { loop(i); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> l(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  int size() const { return static_cast<int>(th_.size()) + 1; }

  void run(int64_t n, int64_t grain, const std::function<void(int64_t, int64_t, int)>& f) {
    if (n <= 0) return;
    if (th_.empty() || n <= grain) {
      f(0, n, 0);
      return;
    }
    {
      std::lock_guard<std::mutex> l(m_);
      fn_ = &f;
      n_ = n;
      grain_ = grain;
      next_.store(0);
      busy_ = static_cast<int>(th_.size());
      ++gen_;
    }
    cv_.notify_all();
    work(0);
    std::unique_lock<std::mutex> l(m_);
    done_cv_.wait(l, [this] { return busy_ == 0; });
    fn_ = nullptr;
  }

 private:
  void work(int id) {
    for (;;) {
      const int64_t b = next_.fetch_add(grain_);
      if (b >= n_) return;
      (*fn_)(b, std::min(n_, b + grain_), id);
    }
  }
  void loop(int id) {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      work(id);
      {
        std::lock_guard<std::mutex> l(m_);
        if (--busy_ == 0) done_cv_.notify_one();
      }
    }
  }

  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int64_t, int64_t, int)>* fn_ = nullptr;
  int64_t n_ = 0, grain_ = 1;
  std::atomic<int64_t> next_{0};
  int busy_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// ------------------------------------------------------ stretch projection

// Closest matrix to the 3x2 deformation gradient F with singular values
// clamped to [lo, hi] (cloth_solver_utils.cpp:225-294: JacobiSVD, clamp,
// U diag(s) V^T), from the eigen decomposition of F^T F.
void project_stretch(const double F[3][2], double lo, double hi, double T[3][2]) {
  const double a = F[0][0] * F[0][0] + F[1][0] * F[1][0] + F[2][0] * F[2][0];
  const double b = F[0][0] * F[0][1] + F[1][0] * F[1][1] + F[2][0] * F[2][1];
  const double c = F[0][1] * F[0][1] + F[1][1] * F[1][1] + F[2][1] * F[2][1];
  const double mean = 0.5 * (a + c);
  const double r = std::sqrt(0.25 * (a - c) * (a - c) + b * b);
  const double l1 = mean + r, l2 = std::max(0.0, mean - r);
  // Eigenvector of l1 (the other is its perpendicular).
  double v1[2];
  if (r == 0.0) {
    v1[0] = 1.0;
    v1[1] = 0.0;
  } else if (a >= c) {
    const double x = a - l2, y = b, n = std::hypot(x, y);
    v1[0] = x / n;
    v1[1] = y / n;
  } else {
    const double x = b, y = c - l2, n = std::hypot(x, y);
    v1[0] = x / n;
    v1[1] = y / n;
  }
  const double v2[2] = {-v1[1], v1[0]};
  const double s1 = std::sqrt(l1), s2 = std::sqrt(l2);
  double u1[3], u2[3];
  for (int i = 0; i < 3; ++i) {
    u1[i] = F[i][0] * v1[0] + F[i][1] * v1[1];
    u2[i] = F[i][0] * v2[0] + F[i][1] * v2[1];
  }
  if (s1 > 0.0) {
    for (double& x : u1) x /= s1;
  } else {
    u1[0] = 1.0;
    u1[1] = u1[2] = 0.0;
  }
  if (s2 > 1e-12 * std::max(1.0, s1)) {
    for (double& x : u2) x /= s2;
  } else {
    // Collapsed face: any unit vector perpendicular to u1.
    const int k = std::fabs(u1[0]) < 0.6 ? 0 : (std::fabs(u1[1]) < 0.6 ? 1 : 2);
    double e[3] = {0.0, 0.0, 0.0};
    e[k] = 1.0;
    const double d = e[0] * u1[0] + e[1] * u1[1] + e[2] * u1[2];
    for (int i = 0; i < 3; ++i) u2[i] = e[i] - d * u1[i];
    const double n = std::sqrt(u2[0] * u2[0] + u2[1] * u2[1] + u2[2] * u2[2]);
    for (double& x : u2) x /= n;
  }
  const double c1 = std::clamp(s1, lo, hi), c2 = std::clamp(s2, lo, hi);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 2; ++j) T[i][j] = c1 * u1[i] * v1[j] + c2 * u2[i] * v2[j];
}

// ------------------------------------------------------------- collision

struct Box {
  float lo[3], hi[3];
};

inline bool overlap(const Box& a, const Box& b) {
  for (int d = 0; d < 3; ++d)
    if (a.hi[d] < b.lo[d] || b.hi[d] < a.lo[d]) return false;
  return true;
}

struct Object {
  bool dynamic = false;
  int cloth = -1;  // index among cloths
  int group = -1;  // -1: collision off
  bool self_on = false;
  float pad = 0.0f;
  int nv = 0;
  const TriMesh* mesh = nullptr;
  std::vector<float> inv_mass;  // per vertex; 0 for pinned and obstacles
  // obstacle frame
  std::vector<float> prev, curr;
};

// A candidate: primitive a of object oa against primitive b of object ob;
// PT pairs keep the triangle first.
struct Cand {
  int oa, a, ob, b;
  bool pt;
};

struct Contact {
  int obj[4], vtx[4];
  float inv_mass[4];
  float corr[4][3];
};

// Positions of one object at the two ends of a segment (3 per vertex).
struct Frame {
  std::vector<const float*> t0, t1;
};

// Barycentrics of the projection of x0 onto triangle (x1, x2, x3), or false
// when the projection falls outside it or the triangle is degenerate
// (cpu/collision/dcd.cpp:7-38, in FP32).
bool pt_uv(const float* x0, const float* x1, const float* x2, const float* x3, float eps,
           float* u, float* v) {
  float e1[3], e2[3], r[3];
  for (int d = 0; d < 3; ++d) {
    e1[d] = x2[d] - x1[d];
    e2[d] = x3[d] - x1[d];
    r[d] = x0[d] - x1[d];
  }
  auto dot = [](const float* p, const float* q) { return p[0] * q[0] + p[1] * q[1] + p[2] * q[2]; };
  const float a = dot(e1, e1), b = dot(e1, e2), c = dot(e2, e2);
  const float pr = dot(e1, r), qr = dot(e2, r);
  const float det = a * c - b * b;
  const float ae = eps * std::max(a, c);
  if (det < ae * ae) return false;
  const float s = (c * pr - b * qr) / det;
  const float t = (a * qr - b * pr) / det;
  if (s < 0.0f || t < 0.0f || s + t > 1.0f) return false;
  *u = s;
  *v = t;
  return true;
}

// Closest points of segments (x0, x1) and (x2, x3) in double, clamped to the
// segments (cpu/collision/dcd.cpp:40-84).
bool ee_uv(const float* x0, const float* x1, const float* x2, const float* x3, float eps,
           float* uo, float* vo) {
  double d1[3], d2[3], r[3];
  for (int d = 0; d < 3; ++d) {
    d1[d] = static_cast<double>(x1[d] - x0[d]);
    d2[d] = static_cast<double>(x3[d] - x2[d]);
    r[d] = static_cast<double>(x0[d] - x2[d]);
  }
  auto dot = [](const double* p, const double* q) { return p[0] * q[0] + p[1] * q[1] + p[2] * q[2]; };
  const double a = dot(d1, d1), b = dot(d1, d2), c = dot(d1, r), e = dot(d2, d2),
               f = dot(d2, r);
  const double e2 = static_cast<double>(eps) * eps;
  const double scale = std::max(a, e);
  if (a <= e2 * scale || e <= e2 * scale) return false;
  const double den = a * e - b * b;
  double u = den > e2 * a * e ? std::clamp((b * f - c * e) / den, 0.0, 1.0) : 0.0;
  double v = (b * u + f) / e;
  if (v < 0.0) {
    v = 0.0;
    u = std::clamp(-c / a, 0.0, 1.0);
  } else if (v > 1.0) {
    v = 1.0;
    u = std::clamp((b - c) / a, 0.0, 1.0);
  }
  *uo = static_cast<float>(u);
  *vo = static_cast<float>(v);
  return true;
}

// ---- interval root finder (FP64) of the CCD of
// cpu/collision/interval_root_finder.cpp:391-510 and ccd.cpp:139-257: the
// distance function over (t, u, v) boxes, evaluated at box corners, boxes
// popped by smallest t.lower (ties: the newest), split along the dimension
// whose width exceeds its tolerance the most.

struct RBox {
  double lo[3], hi[3];
  int64_t seq;
};

struct RBoxLater {
  bool operator()(const RBox& a, const RBox& b) const {
    if (a.lo[0] != b.lo[0]) return a.lo[0] > b.lo[0];
    return a.seq < b.seq;
  }
};

struct Query {
  bool vf = false;
  double p0[4][3], p1[4][3];
  double c[6][3];  // affine coefficients of the distance function per axis
};

void query_coefficients(Query& q) {
  for (int d = 0; d < 3; ++d) {
    const double(*a)[3] = q.p0;
    const double(*b)[3] = q.p1;
    if (q.vf) {
      q.c[0][d] = a[0][d] - a[1][d];
      q.c[1][d] = b[0][d] - b[1][d];
      q.c[2][d] = a[1][d] - a[2][d];
      q.c[3][d] = b[1][d] - b[2][d];
      q.c[4][d] = a[1][d] - a[3][d];
      q.c[5][d] = b[1][d] - b[3][d];
    } else {
      q.c[0][d] = a[0][d] - a[2][d];
      q.c[1][d] = b[0][d] - b[2][d];
      q.c[2][d] = a[1][d] - a[0][d];
      q.c[3][d] = b[1][d] - b[0][d];
      q.c[4][d] = a[2][d] - a[3][d];
      q.c[5][d] = b[2][d] - b[3][d];
    }
  }
}

// Per-dimension tolerances (ccd.cpp:139-193): the requested distance
// accuracy over three times the largest corner displacement along t, u, v.
void dim_tolerance(const Query& q, double dist, double tol[3]) {
  double g[2][4][3];  // [time][corner 000, 001, 011, 010][axis]
  for (int t = 0; t < 2; ++t) {
    const double(*p)[3] = t == 0 ? q.p0 : q.p1;
    for (int d = 0; d < 3; ++d) {
      if (q.vf) {
        g[t][0][d] = p[0][d] - p[1][d];
        g[t][1][d] = p[0][d] - p[3][d];
        g[t][2][d] = p[0][d] - (p[2][d] + p[3][d] - p[1][d]);
        g[t][3][d] = p[0][d] - p[2][d];
      } else {
        g[t][0][d] = p[0][d] - p[2][d];
        g[t][1][d] = p[0][d] - p[3][d];
        g[t][2][d] = p[1][d] - p[3][d];
        g[t][3][d] = p[1][d] - p[2][d];
      }
    }
  }
  auto gap = [&](int t0, int k0, int t1, int k1) {
    double m = 0.0;
    for (int d = 0; d < 3; ++d) m = std::max(m, std::fabs(g[t1][k1][d] - g[t0][k0][d]));
    return m;
  };
  // Along t: same corner at both times; along u: 000-010, 001-011 at both
  // times; along v: 000-001, 010-011 at both times.
  const double dt = 3 * std::max({gap(0, 0, 1, 0), gap(0, 1, 1, 1), gap(0, 2, 1, 2), gap(0, 3, 1, 3)});
  const double du = 3 * std::max({gap(0, 0, 0, 3), gap(1, 0, 1, 3), gap(1, 1, 1, 2), gap(0, 1, 0, 2)});
  const double dv = 3 * std::max({gap(0, 0, 0, 1), gap(1, 0, 1, 1), gap(1, 3, 1, 2), gap(0, 3, 0, 2)});
  auto div = [&](double s) { return s == 0.0 ? std::numeric_limits<double>::infinity() : dist / s; };
  tol[0] = div(dt);
  tol[1] = div(du);
  tol[2] = div(dv);
}

// Range of axis d of the distance function over box b.
void box_range(const Query& q, const RBox& b, int d, double* mn, double* mx) {
  const double* c = nullptr;
  double cc[6];
  for (int i = 0; i < 6; ++i) cc[i] = q.c[i][d];
  c = cc;
  const double ulo = b.lo[1], uhi = b.hi[1], vlo = b.lo[2], vhi = b.hi[2];
  double lo = std::numeric_limits<double>::infinity(), hi = -lo;
  for (int e = 0; e < 2; ++e) {
    const double t = e == 0 ? b.lo[0] : b.hi[0];
    const double q0 = c[0] + (c[1] - c[0]) * t;
    const double q1 = c[2] + (c[3] - c[2]) * t;
    const double q2 = c[4] + (c[5] - c[4]) * t;
    auto add = [&](double u, double v) {
      const double val = q0 + q1 * u + q2 * v;
      lo = std::min(lo, val);
      hi = std::max(hi, val);
    };
    if (!q.vf || uhi + vhi <= 1.0) {
      add(ulo, vlo);
      add(ulo, vhi);
      add(uhi, vlo);
      add(uhi, vhi);
    } else {
      // Rectangle clipped by u + v <= 1: its feasible corners and the
      // points where the diagonal crosses its edges.
      if (ulo + vlo <= 1.0) add(ulo, vlo);
      if (ulo + vhi <= 1.0) add(ulo, vhi);
      if (uhi + vlo <= 1.0) add(uhi, vlo);
      if (1.0 - ulo >= vlo && 1.0 - ulo <= vhi) add(ulo, 1.0 - ulo);
      if (1.0 - uhi >= vlo && 1.0 - uhi <= vhi) add(uhi, 1.0 - uhi);
      if (1.0 - vlo >= ulo && 1.0 - vlo <= uhi) add(1.0 - vlo, vlo);
      if (1.0 - vhi >= ulo && 1.0 - vhi <= uhi) add(1.0 - vhi, vhi);
    }
  }
  *mn = lo;
  *mx = hi;
}

// Earliest t.lower of a box that may hold a root with separation ms, or
// false when there is none before max_t.
bool root_find(const Query& q, const double tol[3], double co_tol, const double err[3],
               double ms, int max_iter, double max_t, double* t_out) {
  std::priority_queue<RBox, std::vector<RBox>, RBoxLater> heap;
  int64_t seq = 0;
  heap.push(RBox{{0.0, 0.0, 0.0}, {max_t, 1.0, 1.0}, seq++});
  int iter = 0;
  while (!heap.empty()) {
    const RBox b = heap.top();
    heap.pop();
    ++iter;
    bool inside_all = true, maybe = true;
    double width[3];
    for (int d = 0; d < 3 && maybe; ++d) {
      double mn, mx;
      box_range(q, b, d, &mn, &mx);
      const double eps = err[d] + ms;
      if (mn > eps || mx < -eps) maybe = false;
      width[d] = mx - mn;
      inside_all = inside_all && mn >= -eps && mx <= eps;
    }
    if (!maybe) continue;
    const bool small = inside_all ||
                       (width[0] <= co_tol && width[1] <= co_tol && width[2] <= co_tol);
    if (small || (max_iter > 0 && iter > max_iter)) {
      *t_out = b.lo[0];
      return true;
    }
    int split = 0;
    double best = -std::numeric_limits<double>::infinity();
    for (int d = 0; d < 3; ++d) {
      const double w = b.hi[d] - b.lo[d];
      const double s = w > tol[d] ? w / tol[d] : -std::numeric_limits<double>::infinity();
      if (s > best) {
        best = s;
        split = d;
      }
    }
    const double mid = 0.5 * (b.lo[split] + b.hi[split]);
    if (mid < kSplitLowerBound) {
      *t_out = b.lo[0];
      return true;
    }
    RBox hi_half = b, lo_half = b;
    hi_half.lo[split] = mid;
    lo_half.hi[split] = mid;
    const bool tri = q.vf && split != 0;
    const double other = split == 1 ? b.lo[2] : b.lo[1];
    if (!tri || hi_half.lo[split] + other <= 1.0) {
      hi_half.seq = seq++;
      heap.push(hi_half);
    }
    if (!tri || lo_half.lo[split] + other <= 1.0) {
      lo_half.seq = seq++;
      heap.push(lo_half);
    }
  }
  return false;
}

// Time of impact of one query within [0, max_t], or false.  A root at t = 0
// is searched again at half the separation; at t = 0 again the pair gives no
// collision for the segment (ccd.cpp:150-180).
bool ccd_query(const Query& q, double ms, double tolerance, int max_iter, const double err[3],
               double max_t, double* toi) {
  double tol[3];
  dim_tolerance(q, tolerance, tol);
  double t;
  if (!root_find(q, tol, tolerance, err, ms, max_iter, max_t, &t)) return false;
  if (t > 0.0) {
    *toi = t;
    return true;
  }
  if (!root_find(q, tol, tolerance, err, 0.5 * ms, max_iter, max_t, &t)) {
    *toi = max_t;
    return true;
  }
  if (t > 0.0) {
    *toi = t;
    return true;
  }
  return false;
}

// Uniform-grid pair search between box sets A and B: every B box is entered
// in the cells it covers; an A box visits its cells and reports a pair once,
// in the cell holding the lower corner of the two boxes' intersection.
class Grid {
 public:
  void build(const std::vector<Box>& B, float cell) {
    h_ = cell;
    entries_.clear();
    for (int b = 0; b < static_cast<int>(B.size()); ++b) {
      int lo[3], hi[3];
      cells(B[b], lo, hi);
      for (int x = lo[0]; x <= hi[0]; ++x)
        for (int y = lo[1]; y <= hi[1]; ++y)
          for (int z = lo[2]; z <= hi[2]; ++z) entries_.emplace_back(key(x, y, z), b);
    }
    std::sort(entries_.begin(), entries_.end());
  }
  template <class Emit>
  void query(const Box& a, const std::vector<Box>& B, Emit&& emit) const {
    int lo[3], hi[3];
    cells(a, lo, hi);
    for (int x = lo[0]; x <= hi[0]; ++x)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int z = lo[2]; z <= hi[2]; ++z) {
          const uint64_t k = key(x, y, z);
          auto it = std::lower_bound(entries_.begin(), entries_.end(),
                                     std::make_pair(k, std::numeric_limits<int>::min()));
          for (; it != entries_.end() && it->first == k; ++it) {
            const Box& b = B[it->second];
            if (!overlap(a, b)) continue;
            const int ox = cell_of(std::max(a.lo[0], b.lo[0]));
            const int oy = cell_of(std::max(a.lo[1], b.lo[1]));
            const int oz = cell_of(std::max(a.lo[2], b.lo[2]));
            if (ox == x && oy == y && oz == z) emit(it->second);
          }
        }
  }

 private:
  int cell_of(float c) const { return static_cast<int>(std::floor(c / h_)); }
  void cells(const Box& b, int lo[3], int hi[3]) const {
    for (int d = 0; d < 3; ++d) {
      lo[d] = cell_of(b.lo[d]);
      hi[d] = cell_of(b.hi[d]);
    }
  }
  static uint64_t key(int x, int y, int z) {
    auto u = [](int v) { return static_cast<uint64_t>(static_cast<uint32_t>(v) & 0x1fffffu); };
    return (u(x) << 42) | (u(y) << 21) | u(z);
  }
  float h_ = 1.0f;
  std::vector<std::pair<uint64_t, int>> entries_;
};

}  // namespace

// ------------------------------------------------------------------ world

struct Cloth {
  const TriMesh* mesh = nullptr;
  silk_hip_cloth_params params{};
  std::vector<int> pins;
  std::vector<float> pin_pos;
  std::vector<double> mass;       // density * Voronoi area
  std::vector<double> face_w;     // k * area per face
  std::vector<float> face_param;  // ia, ib, ic, sqrt(area)
  CsrMatrix H;                    // M/h^2 + k JWJ + kb CWC + pin penalty
  std::vector<double> rest_rhs;   // kb CWC V + penalty * pin targets
  SparseCholesky chol;
  std::vector<double> barrier;    // barrier diagonal the factor holds
  std::vector<float> x, v;
  int off = 0;                    // first vertex in the global state
};

struct CpuWorld::Impl {
  silk_hip_global g{};
  int ccd_max_iter = 4096;
  float ccd_tol = 1e-6f, ms_scale = 0.2f, dcd_scale = 1.0f;
  std::unique_ptr<Pool> pool;
  std::vector<Cloth> cloths;
  std::vector<Object> objs;  // cloths (in order), then obstacles
  int n_state = 0;           // cloth vertices in the global state

  // Positions of object o's vertex at the two ends of a segment.
  Frame frame(const std::vector<float>& a, const std::vector<float>& b) const {
    Frame f;
    for (const Object& o : objs) {
      if (o.dynamic) {
        const int off = cloths[o.cloth].off;
        f.t0.push_back(a.data() + 3 * off);
        f.t1.push_back(b.data() + 3 * off);
      } else {
        f.t0.push_back(o.prev.data());
        f.t1.push_back(o.curr.data());
      }
    }
    return f;
  }

  std::vector<Cand> candidates(const Frame& fr) const;
  bool dcd(const Frame& fr, const Cand& c, float time, float dhat, Contact* out) const;
  bool step(StepStats* st);
};

CpuWorld::CpuWorld(const silk_hip_global& g, int threads) : p_(std::make_unique<Impl>()) {
  p_->g = g;
  if (g.ccd_max_iter > 0) p_->ccd_max_iter = g.ccd_max_iter;
  if (g.ccd_tolerance > 0) p_->ccd_tol = g.ccd_tolerance;
  if (g.ccd_min_separation_scale > 0) p_->ms_scale = g.ccd_min_separation_scale;
  if (g.dcd_activation_scale > 0) p_->dcd_scale = g.dcd_activation_scale;
  if (threads <= 0) {
    // The cores this process may run on (its affinity mask), not the
    // machine's total.
    cpu_set_t set;
    CPU_ZERO(&set);
    threads = sched_getaffinity(0, sizeof(set), &set) == 0
                  ? CPU_COUNT(&set)
                  : static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    threads = std::clamp(threads, 1, 64);
  }
  p_->pool = std::make_unique<Pool>(threads);
}

CpuWorld::~CpuWorld() = default;

int CpuWorld::threads() const { return p_->pool->size(); }

void CpuWorld::add_cloth(const ClothInput& in) {
  Impl& P = *p_;
  Cloth c;
  c.mesh = in.mesh;
  c.params = in.cloth;
  c.pins = in.pins;
  c.pin_pos = in.pin_pos;
  c.x = in.x;
  c.v = in.v;
  const int nv = in.mesh->nv();
  const ClothGeometry geo = build_cloth_geometry(*in.mesh);
  const double h = P.g.dt;
  c.mass.resize(nv);
  for (int i = 0; i < nv; ++i) c.mass[i] = in.cloth.density * geo.mass[i];
  c.face_param = geo.face_param;
  c.face_w.resize(geo.area.size());
  for (size_t f = 0; f < geo.area.size(); ++f) c.face_w[f] = in.cloth.elastic_stiffness * geo.area[f];
  c.H = combine_operator(geo, in.cloth.elastic_stiffness, in.cloth.bending_stiffness);
  std::vector<double> diag(nv, 0.0);
  for (int i = 0; i < nv; ++i) diag[i] = c.mass[i] / (h * h);
  for (int p : c.pins) diag[p] += kPinPenalty;
  for (int i = 0; i < nv; ++i)
    for (int q = c.H.row_ptr[i]; q < c.H.row_ptr[i + 1]; ++q)
      if (c.H.col[q] == i) c.H.val[q] += diag[i];
  c.rest_rhs.resize(3 * static_cast<size_t>(nv));
  for (size_t q = 0; q < c.rest_rhs.size(); ++q)
    c.rest_rhs[q] = in.cloth.bending_stiffness * geo.bending_rest[q];
  c.chol.analyse(c.H, in.mesh->V);
  c.barrier.assign(nv, 0.0);
  if (!c.chol.factor(c.H, c.barrier))
    throw std::runtime_error("CPU backend: system matrix is not positive definite");
  c.off = P.n_state;
  P.n_state += nv;

  Object o;
  o.dynamic = true;
  o.cloth = static_cast<int>(P.cloths.size());
  o.group = in.coll.is_collision_on ? in.coll.group : -1;
  o.self_on = in.coll.is_self_collision_on != 0;
  o.pad = 0.05f * in.mesh->avg_edge_length;
  o.nv = nv;
  o.mesh = in.mesh;
  o.inv_mass.resize(nv);
  for (int i = 0; i < nv; ++i) o.inv_mass[i] = static_cast<float>(1.0 / c.mass[i]);
  for (int p : c.pins) o.inv_mass[p] = 0.0f;
  // Cloths stay before obstacles in the object list.
  int at = 0;
  while (at < static_cast<int>(P.objs.size()) && P.objs[at].dynamic) ++at;
  P.objs.insert(P.objs.begin() + at, std::move(o));
  P.cloths.push_back(std::move(c));
}

void CpuWorld::add_obstacle(const ObstacleInput& in) {
  Object o;
  o.dynamic = false;
  o.group = in.coll.is_collision_on ? in.coll.group : -1;
  o.pad = 0.05f * in.mesh->avg_edge_length;
  o.nv = in.mesh->nv();
  o.mesh = in.mesh;
  o.inv_mass.assign(o.nv, 0.0f);
  o.prev = in.mesh->V;
  o.curr = in.mesh->V;
  p_->objs.push_back(std::move(o));
}

void CpuWorld::set_obstacle_frame(int idx, const float* prev, const float* curr) {
  int k = -1;
  for (Object& o : p_->objs) {
    if (o.dynamic || ++k != idx) continue;
    o.prev.assign(prev, prev + 3 * o.nv);
    o.curr.assign(curr, curr + 3 * o.nv);
  }
}

void CpuWorld::set_pin_targets(int cloth, const float* pos) {
  Cloth& c = p_->cloths.at(cloth);
  c.pin_pos.assign(pos, pos + 3 * c.pins.size());
}

const std::vector<float>& CpuWorld::x(int cloth) const { return p_->cloths.at(cloth).x; }
const std::vector<float>& CpuWorld::v(int cloth) const { return p_->cloths.at(cloth).v; }

bool CpuWorld::step(StepStats* st) { return p_->step(st); }

// Candidate pairs (collision/pipeline.cpp:23-158): object pairs of one group
// with a dynamic member -- triangles of each against the other's points, and
// edges against edges -- then self pairs of cloths with self collision on
// (no shared vertex); pairs whose vertices all have zero inverse mass are
// dropped.  Boxes span both segment ends, padded by the object's padding.
std::vector<Cand> CpuWorld::Impl::candidates(const Frame& fr) const {
  const int no = static_cast<int>(objs.size());
  auto prim_box = [&](int o, const int* v, int k) {
    Box b;
    const float pad = objs[o].pad;
    for (int d = 0; d < 3; ++d) {
      b.lo[d] = std::numeric_limits<float>::max();
      b.hi[d] = -std::numeric_limits<float>::max();
    }
    for (int i = 0; i < k; ++i)
      for (int t = 0; t < 2; ++t) {
        const float* p = (t == 0 ? fr.t0[o] : fr.t1[o]) + 3 * v[i];
        for (int d = 0; d < 3; ++d) {
          b.lo[d] = std::min(b.lo[d], p[d]);
          b.hi[d] = std::max(b.hi[d], p[d]);
        }
      }
    for (int d = 0; d < 3; ++d) {
      b.lo[d] -= pad;
      b.hi[d] += pad;
    }
    return b;
  };
  struct Sets {
    std::vector<Box> pts, edges, tris;
    Box all;
  };
  std::vector<Sets> S(no);
  pool->run(no, 1, [&](int64_t b, int64_t e, int) {
    for (int64_t o = b; o < e; ++o) {
      const Object& ob = objs[o];
      if (ob.group == -1) continue;
      const TriMesh& m = *ob.mesh;
      Sets& s = S[o];
      s.pts.resize(ob.nv);
      for (int i = 0; i < ob.nv; ++i) s.pts[i] = prim_box(static_cast<int>(o), &i, 1);
      s.edges.resize(m.ne());
      for (int i = 0; i < m.ne(); ++i) s.edges[i] = prim_box(static_cast<int>(o), &m.E[2 * i], 2);
      s.tris.resize(m.nf());
      for (int i = 0; i < m.nf(); ++i) s.tris[i] = prim_box(static_cast<int>(o), &m.F[3 * i], 3);
      s.all = s.pts.empty() ? Box{} : s.pts[0];
      for (const Box& x : s.pts)
        for (int d = 0; d < 3; ++d) {
          s.all.lo[d] = std::min(s.all.lo[d], x.lo[d]);
          s.all.hi[d] = std::max(s.all.hi[d], x.hi[d]);
        }
    }
  });
  auto zero_mass = [&](int o, const int* v, int k, int o2, const int* w, int k2) {
    float s = 0.0f;
    for (int i = 0; i < k; ++i) s += objs[o].inv_mass[v[i]];
    for (int i = 0; i < k2; ++i) s += objs[o2].inv_mass[w[i]];
    return s == 0.0f;
  };
  auto mean_extent = [](const std::vector<Box>& B) {
    double s = 0.0;
    for (const Box& b : B)
      s += std::max({b.hi[0] - b.lo[0], b.hi[1] - b.lo[1], b.hi[2] - b.lo[2]});
    return B.empty() ? 1.0f : static_cast<float>(std::max(1e-6, 1.5 * s / B.size()));
  };
  std::vector<Cand> out;
  const int T = pool->size();
  // One pass: query boxes A (object oa, kind ka) against B (object ob, kind kb).
  auto pass = [&](int oa, int ka, int ob, int kb, bool self) {
    const TriMesh& ma = *objs[oa].mesh;
    const TriMesh& mb = *objs[ob].mesh;
    auto set = [&](int o, int k) -> const std::vector<Box>& {
      return k == 0 ? S[o].pts : (k == 1 ? S[o].edges : S[o].tris);
    };
    auto verts = [&](const TriMesh& m, int k, int i, int* n) -> const int* {
      static thread_local int one;
      if (k == 0) {
        one = i;
        *n = 1;
        return &one;
      }
      *n = k == 1 ? 2 : 3;
      return k == 1 ? &m.E[2 * i] : &m.F[3 * i];
    };
    const std::vector<Box>& A = set(oa, ka);
    const std::vector<Box>& B = set(ob, kb);
    if (A.empty() || B.empty()) return;
    Grid grid;
    grid.build(B, std::max(mean_extent(A), mean_extent(B)));
    std::vector<std::vector<Cand>> part(T);
    pool->run(static_cast<int64_t>(A.size()), 256, [&](int64_t b, int64_t e, int t) {
      for (int64_t i = b; i < e; ++i) {
        int na, nb;
        const int* va = verts(ma, ka, static_cast<int>(i), &na);
        int vcopy[3];
        std::copy(va, va + na, vcopy);
        grid.query(A[i], B, [&](int j) {
          if (self && ka == kb && j <= i) return;
          const int* vb = verts(mb, kb, j, &nb);
          if (self)
            for (int x = 0; x < na; ++x)
              for (int y = 0; y < nb; ++y)
                if (vcopy[x] == vb[y]) return;
          if (zero_mass(oa, vcopy, na, ob, vb, nb)) return;
          // PT pairs keep the triangle first.
          if (ka == 0) part[t].push_back(Cand{ob, j, oa, static_cast<int>(i), true});
          else if (kb == 0) part[t].push_back(Cand{oa, static_cast<int>(i), ob, j, true});
          else part[t].push_back(Cand{oa, static_cast<int>(i), ob, j, false});
        });
      }
    });
    for (auto& p : part) out.insert(out.end(), p.begin(), p.end());
  };
  for (int a = 0; a < no; ++a)
    for (int b = a + 1; b < no; ++b) {
      const Object &A = objs[a], &B = objs[b];
      if (A.group == -1 || B.group == -1 || A.group != B.group) continue;
      if (!A.dynamic && !B.dynamic) continue;
      if (!overlap(S[a].all, S[b].all)) continue;
      pass(a, 0, b, 2, false);  // points of a, triangles of b
      pass(a, 2, b, 0, false);  // triangles of a, points of b
      pass(a, 1, b, 1, false);  // edges
    }
  for (int a = 0; a < no; ++a) {
    const Object& A = objs[a];
    if (!A.dynamic || !A.self_on || A.group == -1) continue;
    pass(a, 0, a, 2, true);
    pass(a, 1, a, 1, true);
  }
  return out;
}

// DCD of one candidate at frame time `time` (narrowphase.cpp:124-218): a
// contact when the closest points are nearer than dhat, with the correction
// (dhat - d) n distributed by barycentric weight over inverse mass.
bool CpuWorld::Impl::dcd(const Frame& fr, const Cand& c, float time, float dhat,
                         Contact* out) const {
  auto pos = [&](int o, int v, float* p) {
    const float* a = fr.t0[o] + 3 * v;
    const float* b = fr.t1[o] + 3 * v;
    for (int d = 0; d < 3; ++d) p[d] = objs[o].dynamic ? b[d] : a[d] + time * (b[d] - a[d]);
  };
  float x[4][3], w[4];
  int obj[4], vtx[4];
  if (c.pt) {
    const TriMesh& mt = *objs[c.oa].mesh;
    obj[0] = c.ob;
    vtx[0] = c.b;
    for (int q = 0; q < 3; ++q) {
      obj[1 + q] = c.oa;
      vtx[1 + q] = mt.F[3 * c.a + q];
    }
  } else {
    const TriMesh& ma = *objs[c.oa].mesh;
    const TriMesh& mb = *objs[c.ob].mesh;
    for (int q = 0; q < 2; ++q) {
      obj[q] = c.oa;
      vtx[q] = ma.E[2 * c.a + q];
      obj[2 + q] = c.ob;
      vtx[2 + q] = mb.E[2 * c.b + q];
    }
  }
  for (int q = 0; q < 4; ++q) pos(obj[q], vtx[q], x[q]);
  float pa[3], pb[3];
  if (c.pt) {
    float u, v;
    if (!pt_uv(x[0], x[1], x[2], x[3], kDcdEps, &u, &v)) return false;
    w[0] = 1.0f;
    w[1] = 1.0f - u - v;
    w[2] = u;
    w[3] = v;
    for (int d = 0; d < 3; ++d) {
      pa[d] = x[0][d];
      pb[d] = w[1] * x[1][d] + w[2] * x[2][d] + w[3] * x[3][d];
    }
  } else {
    float u, v;
    if (!ee_uv(x[0], x[1], x[2], x[3], kDcdEps, &u, &v)) return false;
    w[0] = 1.0f - u;
    w[1] = u;
    w[2] = 1.0f - v;
    w[3] = v;
    for (int d = 0; d < 3; ++d) {
      pa[d] = w[0] * x[0][d] + w[1] * x[1][d];
      pb[d] = w[2] * x[2][d] + w[3] * x[3][d];
    }
  }
  float n[3];
  for (int d = 0; d < 3; ++d) n[d] = pb[d] - pa[d];
  const float dist = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (dist >= dhat || dist == 0.0f) return false;
  for (float& a : n) a /= dist;
  float im[4], den = 0.0f;
  for (int q = 0; q < 4; ++q) {
    im[q] = objs[obj[q]].inv_mass[vtx[q]];
    den += w[q] * w[q] * im[q];
  }
  if (den == 0.0f) return false;
  for (int q = 0; q < 4; ++q) {
    float wq = w[q] * im[q] / den;
    if (c.pt ? q == 0 : q < 2) wq = -wq;
    out->obj[q] = obj[q];
    out->vtx[q] = vtx[q];
    out->inv_mass[q] = im[q];
    for (int d = 0; d < 3; ++d) out->corr[q][d] = (dhat - dist) * n[d] * wq;
  }
  return true;
}

bool CpuWorld::Impl::step(StepStats* st) {
  StepStats s;
  const int n3 = 3 * n_state;
  if (n3 == 0) {
    if (st) *st = s;
    return true;
  }
  const float h = g.dt;
  const float acc[3] = {g.acceleration[0], g.acceleration[1], g.acceleration[2]};
  // Global state with damped velocities (pipeline.cpp:186-224).
  std::vector<float> curr(n3), vel(n3);
  for (const Cloth& c : cloths)
    for (size_t q = 0; q < c.x.size(); ++q) {
      curr[3 * c.off + q] = c.x[q];
      vel[3 * c.off + q] = (1.0f - c.params.damping) * c.v[q];
    }
  float bmin[3] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
                   std::numeric_limits<float>::max()};
  float bmax[3] = {-bmin[0], -bmin[1], -bmin[2]};
  for (int q = 0; q < n3; ++q) {
    bmin[q % 3] = std::min(bmin[q % 3], curr[q]);
    bmax[q % 3] = std::max(bmax[q % 3], curr[q]);
  }
  const float scene_scale = std::sqrt((bmax[0] - bmin[0]) * (bmax[0] - bmin[0]) +
                                      (bmax[1] - bmin[1]) * (bmax[1] - bmin[1]) +
                                      (bmax[2] - bmin[2]) * (bmax[2] - bmin[2]));
  // Step-invariant right-hand side: rest curvature and pin targets.
  std::vector<double> init_rhs(n3, 0.0);
  for (const Cloth& c : cloths) {
    for (size_t q = 0; q < c.rest_rhs.size(); ++q) init_rhs[3 * c.off + q] = c.rest_rhs[q];
    for (size_t p = 0; p < c.pins.size(); ++p)
      for (int d = 0; d < 3; ++d)
        init_rhs[3 * (c.off + c.pins[p]) + d] += kPinPenalty * c.pin_pos[3 * p + d];
  }
  float pad = std::numeric_limits<float>::max();
  bool any_coll = false;
  for (const Object& o : objs) {
    pad = std::min(pad, o.pad);
    any_coll = any_coll || o.group != -1;
  }
  const float dhat = dcd_scale * pad, ms = ms_scale * pad;
  const int T = pool->size();

  auto find_contacts = [&](float time) {
    std::vector<Contact> all;
    if (!any_coll) return all;
    const Frame fr = frame(curr, curr);
    const std::vector<Cand> cand = candidates(fr);
    std::vector<std::vector<Contact>> part(T);
    pool->run(static_cast<int64_t>(cand.size()), 512, [&](int64_t b, int64_t e, int t) {
      Contact c;
      for (int64_t i = b; i < e; ++i)
        if (dcd(fr, cand[i], time, dhat, &c)) part[t].push_back(c);
    });
    for (auto& p : part) all.insert(all.end(), p.begin(), p.end());
    return all;
  };

  std::vector<Contact> contacts = find_contacts(0.0f);
  float remaining = 1.0f;
  std::vector<float> next(n3), solution(n3);
  std::vector<double> outer_rhs(n3), blhs(n_state), brhs(n3);
  std::vector<std::vector<double>> part_rhs(T);
  const int max_outer = g.max_outer_iteration > 0 ? g.max_outer_iteration : 100;
  const int max_inner = g.max_inner_iteration > 0 ? g.max_inner_iteration : 100;
  for (int it = 0; it < max_outer; ++it) {
    ++s.outer;
    s.collisions += static_cast<int>(contacts.size());
    // Barrier reflection targets (pipeline.cpp:251-292).
    std::fill(blhs.begin(), blhs.end(), 0.0);
    std::fill(brhs.begin(), brhs.end(), 0.0);
    for (const Contact& c : contacts)
      for (int q = 0; q < 4; ++q) {
        if (c.inv_mass[q] == 0.0f || !objs[c.obj[q]].dynamic) continue;
        const int gv = cloths[objs[c.obj[q]].cloth].off + c.vtx[q];
        blhs[gv] += kBarrierStiffness;
        for (int d = 0; d < 3; ++d)
          brhs[3 * gv + d] += kBarrierStiffness * (curr[3 * gv + d] + c.corr[q][d]);
      }
    for (int q = 0; q < n3; ++q) next[q] = curr[q] + h * vel[q] + h * h * acc[q % 3];
    // Outer right-hand side and the factor with the barrier diagonal
    // (cloth_solver_utils.cpp:117-198).
    for (Cloth& c : cloths) {
      const int nv = static_cast<int>(c.mass.size());
      std::vector<double> bar(nv);
      for (int v = 0; v < nv; ++v) {
        const int gvtx = c.off + v;
        const double m = c.mass[v];
        for (int d = 0; d < 3; ++d) {
          const int q = 3 * gvtx + d;
          outer_rhs[q] = init_rhs[q] + m / (static_cast<double>(h) * h) * curr[q] +
                         m / h * vel[q] + m * acc[d] + brhs[q];
        }
        bar[v] = blhs[gvtx];
      }
      if (bar != c.barrier) {
        if (!c.chol.refactor(c.H, c.barrier, bar)) return false;
        s.rows_refactored += c.chol.rows_recomputed();
        c.barrier = std::move(bar);
      }
    }
    // Inner loop: per-face projections and the global solve until the update
    // is below 5 % of the scene size (pipeline.cpp:96-121).
    for (int in = 0; in < max_inner; ++in) {
      ++s.inner;
      for (Cloth& c : cloths) {
        const TriMesh& m = *c.mesh;
        const int nv = m.nv(), nf = m.nf();
        pool->run(nf, 512, [&](int64_t b, int64_t e, int t) {
          std::vector<double>& rhs = part_rhs[t];
          if (rhs.size() != 3 * static_cast<size_t>(nv)) rhs.assign(3 * static_cast<size_t>(nv), 0.0);
          for (int64_t f = b; f < e; ++f) {
            const int* tri = &m.F[3 * f];
            const float ia = c.face_param[4 * f], ib = c.face_param[4 * f + 1],
                        ic = c.face_param[4 * f + 2];
            const float* x0 = &next[3 * (c.off + tri[0])];
            const float* x1 = &next[3 * (c.off + tri[1])];
            const float* x2 = &next[3 * (c.off + tri[2])];
            double F[3][2], Tm[3][2];
            for (int d = 0; d < 3; ++d) {
              const double e0 = static_cast<double>(x1[d]) - x0[d];
              const double e1 = static_cast<double>(x2[d]) - x0[d];
              F[d][0] = e0 * ia;
              F[d][1] = e0 * ib + e1 * ic;
            }
            project_stretch(F, 0.9, 1.1, Tm);
            // k area G^T applied to T: the constant part of the elastic force.
            const double w = c.face_w[f];
            for (int d = 0; d < 3; ++d) {
              const double t0 = w * Tm[d][0], t1 = w * Tm[d][1];
              rhs[3 * tri[0] + d] -= t0 * ia + t1 * (ib + ic);
              rhs[3 * tri[1] + d] += t0 * ia + t1 * ib;
              rhs[3 * tri[2] + d] += t1 * ic;
            }
          }
        });
        std::vector<double> rhs(outer_rhs.begin() + 3 * c.off,
                                outer_rhs.begin() + 3 * (c.off + nv));
        for (auto& pr : part_rhs) {
          if (pr.size() != rhs.size()) continue;
          for (size_t q = 0; q < rhs.size(); ++q) rhs[q] += pr[q];
          std::fill(pr.begin(), pr.end(), 0.0);
        }
        std::vector<double> xs(rhs.size());
        c.chol.solve3(rhs.data(), xs.data());
        for (size_t q = 0; q < xs.size(); ++q) solution[3 * c.off + q] = static_cast<float>(xs[q]);
      }
      double dn = 0.0;
      for (int q = 0; q < n3; ++q) {
        if (!std::isfinite(solution[q])) return false;
        const double t = static_cast<double>(solution[q]) - next[q];
        dn += t * t;
      }
      next.swap(solution);
      if (std::sqrt(dn) <= 0.05 * scene_scale) break;
    }
    // Barrier dofs to their targets (pipeline.cpp:294-319).
    if (!contacts.empty())
      for (int v = 0; v < n_state; ++v)
        if (blhs[v] != 0.0)
          for (int d = 0; d < 3; ++d)
            next[3 * v + d] = static_cast<float>(brhs[3 * v + d] / blhs[v]);
    // Earliest time of impact over the segment curr -> next
    // (collision/pipeline.cpp:160-210).
    const float time_start = 1.0f - remaining;
    float toi = remaining;
    if (any_coll) {
      const Frame fr = frame(curr, next);
      const std::vector<Cand> cand = candidates(fr);
      s.ccd_candidates += static_cast<int>(cand.size());
      double abs_max[3];
      for (int d = 0; d < 3; ++d) abs_max[d] = std::max(std::fabs(bmin[d]), std::fabs(bmax[d]));
      for (int q = 0; q < n3; ++q)
        abs_max[q % 3] = std::max({abs_max[q % 3], static_cast<double>(std::fabs(curr[q])),
                                   static_cast<double>(std::fabs(next[q]))});
      for (const Object& o : objs)
        if (!o.dynamic)
          for (int q = 0; q < 3 * o.nv; ++q)
            abs_max[q % 3] = std::max({abs_max[q % 3], static_cast<double>(std::fabs(o.prev[q])),
                                       static_cast<double>(std::fabs(o.curr[q]))});
      // Numerical filters of interval_root_finder.cpp:516-517.
      double err_vf[3], err_ee[3];
      for (int d = 0; d < 3; ++d) {
        const double sc = std::max(abs_max[d], 1.0);
        err_vf[d] = static_cast<double>(4.053116e-06f) * sc;
        err_ee[d] = static_cast<double>(3.814698e-06f) * sc;
      }
      std::atomic<float> earliest{remaining};
      pool->run(static_cast<int64_t>(cand.size()), 256, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i) {
          const Cand& c = cand[i];
          Query q;
          q.vf = c.pt;
          auto take = [&](int o, int v, int slot) {
            const float* a = fr.t0[o] + 3 * v;
            const float* bb = fr.t1[o] + 3 * v;
            for (int d = 0; d < 3; ++d) {
              if (objs[o].dynamic) {
                q.p0[slot][d] = a[d];
                q.p1[slot][d] = bb[d];
              } else {
                const float delta = bb[d] - a[d];
                const float p0 = a[d] + time_start * delta;
                q.p0[slot][d] = p0;
                q.p1[slot][d] = p0 + delta;
              }
            }
          };
          if (c.pt) {
            const TriMesh& mt = *objs[c.oa].mesh;
            take(c.ob, c.b, 0);
            for (int k = 0; k < 3; ++k) take(c.oa, mt.F[3 * c.a + k], 1 + k);
          } else {
            const TriMesh& ma = *objs[c.oa].mesh;
            const TriMesh& mb = *objs[c.ob].mesh;
            take(c.oa, ma.E[2 * c.a], 0);
            take(c.oa, ma.E[2 * c.a + 1], 1);
            take(c.ob, mb.E[2 * c.b], 2);
            take(c.ob, mb.E[2 * c.b + 1], 3);
          }
          query_coefficients(q);
          const float max_t = earliest.load(std::memory_order_relaxed);
          double t;
          if (!ccd_query(q, ms, ccd_tol, ccd_max_iter, c.pt ? err_vf : err_ee, max_t, &t))
            continue;
          const float tf = static_cast<float>(t);
          float prev = earliest.load(std::memory_order_relaxed);
          while (tf < prev && !earliest.compare_exchange_weak(prev, tf, std::memory_order_relaxed)) {
          }
        }
      });
      toi = earliest.load();
    }
    for (int q = 0; q < n3; ++q) vel[q] = (next[q] - curr[q]) / h;
    if (toi >= remaining) {
      for (int q = 0; q < n3; ++q) curr[q] += remaining * (next[q] - curr[q]);
      break;
    }
    s.min_toi = std::min(s.min_toi, time_start + toi);
    const float rb = 0.8f * toi;
    for (int q = 0; q < n3; ++q) curr[q] = rb * (next[q] - curr[q]) + curr[q];
    remaining -= rb;
    contacts = find_contacts(1.0f - remaining);
  }
  for (Cloth& c : cloths)
    for (size_t q = 0; q < c.x.size(); ++q) {
      c.x[q] = curr[3 * c.off + q];
      c.v[q] = vel[3 * c.off + q];
    }
  for (Object& o : objs)
    if (!o.dynamic) o.prev = o.curr;
  if (st) *st = s;
  return true;
}

}  // namespace silk::amd::cpu
