// Oracle: the whole GPU-path step (main_loop.cu:174-312) in double.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// Vertex numbering: cloth vertices first (cloth insertion order, each cloth in
// the solver order given by the partition passed to add_cloth), then obstacle
// vertices (insertion order, caller order).  Contacts are keyed, sorted and
// coloured on that numbering so that the GPU library (which uses the same
// rules) visits them in the same order.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <tuple>
#include <vector>

#include "oracle.hpp"

namespace oracle {

struct Params {
  double acc[3] = {0, 0, -10};
  int max_outer = 100, max_inner = 100;
  double dt = 1.0 / 60.0;
  double lin_abs = 1e-7, lin_rel_min = 1e-6, lin_rel_max = 1e-3, lin_rel_init = 1e-2;
  double admm_abs = 1e-5, admm_rel = 1e-3;
  int pcg_max_iter = 1000, sweeps = 10, ccd_max_iter = 4096;
  double ccd_tol = 1e-6, ms_scale = 0.2, dcd_scale = 1.0;
  int pcg_criterion = 0;
  int ccd_zero = 0;  // CCD zero-TOI mode (0: ms / 2 only, 1: then zero separation)
};

struct Coll {
  int on = 1, self_on = 1, group = 0;
  float restitution = 0.3f, friction = 0.3f;
};

struct OCloth {
  double k = 100, kb = 1e-4, density = 0.1, damping = 0.01;
  Coll coll;
  int nv = 0, nf = 0, off = 0;
  std::vector<int> perm, part_offsets;  // caller -> solver
  std::vector<double> V0;               // rest, solver order
  std::vector<int> F, E;                // solver order
  std::vector<char> pinned;
  std::vector<double> pin_target;       // 3*nv
  L2Cache l2;
  Bsr3 A;
  MasOracle mas;
  std::vector<double> mass, cw, bend;   // nv, 3nv, 3nv
  std::vector<double> x, v;             // 3nv
  std::vector<double> z, u;             // 6nf
  float padding = 0.0f;
};

struct OObstacle {
  Coll coll;
  int nv = 0, off = 0;
  std::vector<float> curr, prev;
  std::vector<int> F, E;
  bool is_static = true, is_static_twice = true;
  bool has_pos = false;
  float padding = 0.0f;
};

struct CollisionRec {
  int type = 0;     // 0 PT, 1 EE
  int gv[4];        // global vertex ids
  double coef[4];
  double inv_mass[4];
  double normal[3];
  double x[4][3];   // positions at the DCD time
  double presc[3];  // prescribed obstacle displacement
  double friction = 0, dhat = 0;
};

struct Contact {
  int state[4];  // global dof/3 or -1
  double coef[4], gamma[4];
  double normal[3], affine[3];
  double mu = 0;
  double corr[3] = {0, 0, 0};
};

struct StepStats {
  int outer = 0, admm = 0, pcg = 0, contacts = 0, colors = 0, ccd_cand = 0;
  int ccd_zero = 0;  // candidates left with a root at t = 0 (no collision)
  double min_toi = 1.0;
};

struct World {
  Params P;
  std::vector<OCloth> cloths;
  std::vector<OObstacle> obstacles;
  // per-step
  int nv_cloth = 0, nv_total = 0;
  std::vector<double> pos_t0, pos_t1;  // all vertices (cloth + obstacle), 3 each
  std::vector<Contact> contacts;
  std::set<std::array<int, 5>> keys;
  std::vector<std::vector<int>> color_groups;
  std::vector<double> cbase_w, cw, p, p_prev, eta;  // per dof of cloth state
  int active_dof = 0;
  StepStats stats;
  // Decisions of the last step: CCD TOI of every outer iteration, appended
  // contacts (DCD call, type, four global ids in record order).
  std::vector<double> log_toi;
  std::vector<int> log_admm;  // ADMM iterations of each outer iteration
  std::vector<int> log_info;  // per contact: colour << 4 | mask of its solved vertices
  std::vector<int> log_contacts;
  // Replay of another solver's decisions for the next step (set_replay):
  // the outer loop takes these TOIs and every DCD call appends exactly these
  // contacts, their records formed on the oracle's own positions.
  bool replay = false;
  std::vector<double> rp_toi;
  std::vector<int> rp_admm;  // empty: run the convergence test
  std::vector<int> rp_info;  // empty: colour and mask computed here
  std::vector<int> rp_contacts;
  int replay_missing = 0;  // replayed contacts whose record could not be formed
};

namespace {

int cloth_of(const World& w, int gv) {
  for (size_t c = 0; c < w.cloths.size(); ++c)
    if (gv < w.cloths[c].off + w.cloths[c].nv) return static_cast<int>(c);
  return -1;
}

// Unique sorted edges of a face list, float average length over them.
void edges_and_avg(const std::vector<float>& V, const std::vector<int>& F,
                   std::vector<int>& E, float& avg) {
  std::vector<std::pair<int, int>> e;
  for (size_t f = 0; f < F.size() / 3; ++f)
    for (int k = 0; k < 3; ++k) {
      int a = F[3 * f + k], b = F[3 * f + (k + 1) % 3];
      e.emplace_back(std::min(a, b), std::max(a, b));
    }
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  E.clear();
  float s = 0.0f;
  for (auto& [a, b] : e) {
    E.push_back(a);
    E.push_back(b);
    float dx = V[3 * a] - V[3 * b], dy = V[3 * a + 1] - V[3 * b + 1],
          dz = V[3 * a + 2] - V[3 * b + 2];
    s += std::sqrt(dx * dx + dy * dy + dz * dz);
  }
  avg = e.empty() ? 0.0f : s / static_cast<float>(e.size());
}

// ----------------------------------------------------------- colliders

struct Prim {
  int obj;      // object index: cloth c -> c, obstacle o -> ncloth + o
  int v[3];     // global vertex ids (point: v[0]; edge: v[0..1])
  int nvert;
  Box box;
};

struct ObjInfo {
  bool physical, self_on;
  int group;  // -1 disabled
  float padding;
  float friction;
  Box bbox;
  std::vector<Prim> P, E, T;
};

Box swept_box(const World& w, const int* v, int n, float pad) {
  Box b;
  for (int d = 0; d < 3; ++d) {
    float mn = 1e30f, mx = -1e30f;
    for (int k = 0; k < n; ++k) {
      float a = static_cast<float>(w.pos_t0[3 * v[k] + d]);
      float c = static_cast<float>(w.pos_t1[3 * v[k] + d]);
      mn = std::min(mn, std::min(a, c));
      mx = std::max(mx, std::max(a, c));
    }
    b.mn[d] = mn - pad;
    b.mx[d] = mx + pad;
  }
  return b;
}

double inv_mass_of(const World& w, int gv) {
  int c = cloth_of(w, gv);
  if (c < 0) return 0.0;
  const OCloth& cl = w.cloths[c];
  const int lv = gv - cl.off;
  if (cl.pinned[lv]) return 0.0;
  return 1.0 / cl.mass[lv];
}

std::vector<ObjInfo> build_objects(const World& w) {
  std::vector<ObjInfo> out;
  auto fill = [&](ObjInfo& o, int obj, int off, int nv, const std::vector<int>& F,
                  const std::vector<int>& E) {
    for (int i = 0; i < nv; ++i) {
      Prim p{obj, {off + i, 0, 0}, 1, {}};
      p.box = swept_box(w, p.v, 1, o.padding);
      o.P.push_back(p);
    }
    for (size_t e = 0; e < E.size() / 2; ++e) {
      Prim p{obj, {off + E[2 * e], off + E[2 * e + 1], 0}, 2, {}};
      p.box = swept_box(w, p.v, 2, o.padding);
      o.E.push_back(p);
    }
    for (size_t f = 0; f < F.size() / 3; ++f) {
      Prim p{obj, {off + F[3 * f], off + F[3 * f + 1], off + F[3 * f + 2]}, 3, {}};
      p.box = swept_box(w, p.v, 3, o.padding);
      o.T.push_back(p);
    }
    Box bb{{1e30f, 1e30f, 1e30f}, {-1e30f, -1e30f, -1e30f}};
    for (auto& t : o.T)
      for (int d = 0; d < 3; ++d) {
        bb.mn[d] = std::min(bb.mn[d], t.box.mn[d]);
        bb.mx[d] = std::max(bb.mx[d], t.box.mx[d]);
      }
    o.bbox = bb;
  };
  int obj = 0;
  for (const auto& c : w.cloths) {
    ObjInfo o;
    o.physical = true;
    o.self_on = c.coll.self_on;
    o.group = c.coll.on ? c.coll.group : -1;
    o.padding = c.padding;
    o.friction = c.coll.friction;
    fill(o, obj++, c.off, c.nv, c.F, c.E);
    out.push_back(std::move(o));
  }
  for (const auto& ob : w.obstacles) {
    ObjInfo o;
    o.physical = false;
    o.self_on = ob.coll.self_on;
    o.group = ob.coll.on ? ob.coll.group : -1;
    o.padding = ob.padding;
    o.friction = ob.coll.friction;
    fill(o, obj++, ob.off, ob.nv, ob.F, ob.E);
    out.push_back(std::move(o));
  }
  return out;
}

// Candidate pairs (find_collision.cu:45-166): (triangle, point) and
// (edge, edge) in canonical form.
struct Candidates {
  std::vector<std::pair<const Prim*, const Prim*>> pt;  // (tri, point)
  std::vector<std::pair<const Prim*, const Prim*>> ee;  // (edge a, edge b)
};

Candidates candidates(const World& w, const std::vector<ObjInfo>& objs) {
  Candidates c;
  auto boxes = [](const std::vector<Prim>& v) {
    std::vector<Box> b;
    b.reserve(v.size());
    for (auto& p : v) b.push_back(p.box);
    return b;
  };
  auto all_pinned = [&](const Prim* a, const Prim* b) {
    double s = 0.0;
    for (int k = 0; k < a->nvert; ++k) s += inv_mass_of(w, a->v[k]);
    for (int k = 0; k < b->nvert; ++k) s += inv_mass_of(w, b->v[k]);
    return s == 0.0;
  };
  auto add_pt = [&](const ObjInfo& tobj, const ObjInfo& pobj, bool self) {
    auto pairs = overlap_pairs(boxes(tobj.T), boxes(pobj.P), false);
    for (auto& [ti, pi] : pairs) {
      const Prim* t = &tobj.T[ti];
      const Prim* p = &pobj.P[pi];
      if (all_pinned(t, p)) continue;
      if (self && (t->v[0] == p->v[0] || t->v[1] == p->v[0] || t->v[2] == p->v[0]))
        continue;
      c.pt.emplace_back(t, p);
    }
  };
  auto add_ee = [&](const ObjInfo& a, const ObjInfo& b, bool self) {
    auto pairs = overlap_pairs(boxes(a.E), boxes(self ? a.E : b.E), self);
    for (auto& [i, j] : pairs) {
      const Prim* e1 = &a.E[i];
      const Prim* e2 = self ? &a.E[j] : &b.E[j];
      if (all_pinned(e1, e2)) continue;
      if (self && (e1->v[0] == e2->v[0] || e1->v[0] == e2->v[1] ||
                   e1->v[1] == e2->v[0] || e1->v[1] == e2->v[1]))
        continue;
      // canonical order: smaller (min, max) global edge key first
      auto key = [](const Prim* e) {
        return std::make_pair(std::min(e->v[0], e->v[1]), std::max(e->v[0], e->v[1]));
      };
      if (key(e2) < key(e1)) std::swap(e1, e2);
      c.ee.emplace_back(e1, e2);
    }
  };
  for (size_t i = 0; i < objs.size(); ++i)
    for (size_t j = i + 1; j < objs.size(); ++j) {
      const ObjInfo& a = objs[i];
      const ObjInfo& b = objs[j];
      if (a.group == -1 || b.group == -1 || a.group != b.group) continue;
      if (!(a.physical || b.physical)) continue;
      if (!box_overlap(a.bbox, b.bbox)) continue;
      add_pt(a, b, false);
      add_pt(b, a, false);
      add_ee(a, b, false);
    }
  for (const auto& o : objs) {
    if (!o.physical || !o.self_on) continue;
    add_pt(o, o, true);
    add_ee(o, o, true);
  }
  return c;
}

V3 lerp(const double* a, const double* b, double t) {
  return {(1.0 - t) * a[0] + t * b[0], (1.0 - t) * a[1] + t * b[1],
          (1.0 - t) * a[2] + t * b[2]};
}

}  // namespace

// exact_pt_uv / exact_ee_uv (dcd.cu:321-394).
// clip = false keeps a projection that falls outside the triangle (used
// when replaying a contact that another solver activated).
static bool pt_uv_impl(const V3& x0, const V3& x1, const V3& x2, const V3& x3,
                       double eps, bool clip, double* u, double* v) {
  V3 x21 = x2 - x1, x31 = x3 - x1, x01 = x0 - x1;
  double a = dot(x21, x21), b = dot(x21, x31), c = dot(x31, x31);
  double d = dot(x21, x01), e = dot(x31, x01);
  double det = a * c - b * b;
  double area_eps = eps * std::max(a, c);
  if (det < area_eps * area_eps) return false;
  double b1 = (c * d - b * e) / det, b2 = (a * e - b * d) / det;
  if (clip && (b1 < 0 || b2 < 0 || b1 + b2 > 1)) return false;
  *u = b1;
  *v = b2;
  return true;
}

bool pt_uv(const V3& x0, const V3& x1, const V3& x2, const V3& x3, double eps,
           double* u, double* v) {
  return pt_uv_impl(x0, x1, x2, x3, eps, true, u, v);
}

bool ee_uv(const V3& x0, const V3& x1, const V3& x2, const V3& x3, double eps,
           double* uo, double* vo) {
  V3 d1 = x1 - x0, d2 = x3 - x2, r = x0 - x2;
  double a = dot(d1, d1), b = dot(d1, d2), c = dot(d1, r), e = dot(d2, d2),
         f = dot(d2, r);
  double eps2 = eps * eps, ls2 = std::max(a, e);
  if (a <= eps2 * ls2 || e <= eps2 * ls2) return false;
  double den = a * e - b * b;
  double u = den > eps2 * a * e ? std::clamp((b * f - c * e) / den, 0.0, 1.0) : 0.0;
  double v = (b * u + f) / e;
  if (v < 0.0) {
    v = 0.0;
    u = std::clamp(-c / a, 0.0, 1.0);
  } else if (v > 1.0) {
    v = 1.0;
    u = std::clamp((b - c) / a, 0.0, 1.0);
  }
  *uo = u;
  *vo = v;
  return true;
}


void coulomb_solve(const double us[3], const double n[3], double mu,
                   double out[3]) {
  const double un = us[0] * n[0] + us[1] * n[1] + us[2] * n[2];
  if (un >= 0.0) {
    for (int d = 0; d < 3; ++d) out[d] = us[d];
    return;
  }
  double ut[3];
  for (int d = 0; d < 3; ++d) ut[d] = us[d] - un * n[d];
  const double utn = std::sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
  const double rad = -mu * un;
  if (utn <= rad) {
    out[0] = out[1] = out[2] = 0.0;
    return;
  }
  for (int d = 0; d < 3; ++d) out[d] = (1.0 - rad / utn) * ut[d];
}

// Closest-point record of one PT (gv = point, tri0..2) or EE (gv = a0, a1,
// b0, b1) pair on explicit positions (dcd.cu:102-261): false when the
// closest point lies outside the triangle, the edges are degenerate, the
// distance is 0 or at least dhat, or no vertex can move.  friction and presc
// are left for the caller.
bool dcd_pair_forced(bool pt, const int gv[4], const V3 x[4], const double inv[4],
                     double dhat, bool force, CollisionRec* out);

bool dcd_pair(bool pt, const int gv[4], const V3 x[4], const double inv[4],
              double dhat, CollisionRec* out) {
  return dcd_pair_forced(pt, gv, x, inv, dhat, false, out);
}

bool dcd_pair_forced(bool pt, const int gv[4], const V3 x[4], const double inv[4],
                     double dhat, bool force, CollisionRec* out) {
  const double eps = 1e-6;
  double u, v;
  V3 disp;
  double param[4], coef[4];
  if (pt) {
    if (!pt_uv_impl(x[0], x[1], x[2], x[3], eps, !force, &u, &v)) return false;
    V3 tp = (1.0 - u - v) * x[1] + u * x[2] + v * x[3];
    disp = tp - x[0];
    param[0] = 1.0; param[1] = 1.0 - u - v; param[2] = u; param[3] = v;
    coef[0] = 1.0; coef[1] = -param[1]; coef[2] = -param[2]; coef[3] = -param[3];
  } else {
    if (!ee_uv(x[0], x[1], x[2], x[3], eps, &u, &v)) return false;
    V3 pa = (1.0 - u) * x[0] + u * x[1];
    V3 pb = (1.0 - v) * x[2] + v * x[3];
    disp = pb - pa;
    param[0] = 1.0 - u; param[1] = u; param[2] = 1.0 - v; param[3] = v;
    coef[0] = param[0]; coef[1] = param[1]; coef[2] = -param[2]; coef[3] = -param[3];
  }
  const double d2 = dot(disp, disp);
  if (d2 == 0.0) return false;
  const double dist = std::sqrt(d2);
  if (dist >= dhat && !force) return false;
  V3 n = (-1.0 / dist) * disp;
  double den = 0.0;
  for (int k = 0; k < 4; ++k) den += param[k] * param[k] * inv[k];
  if (den == 0.0) return false;
  CollisionRec r;
  r.type = pt ? 0 : 1;
  for (int k = 0; k < 4; ++k) {
    r.gv[k] = gv[k];
    r.coef[k] = coef[k];
    r.inv_mass[k] = inv[k];
    for (int d = 0; d < 3; ++d) r.x[k][d] = x[k][d];
  }
  for (int d = 0; d < 3; ++d) {
    r.normal[d] = n[d];
    r.presc[d] = 0.0;
  }
  r.dhat = dhat;
  *out = r;
  return true;
}

namespace {

// DCD at fixed time (dcd.cu:102-261).  Cloth colliders are refit with
// t0 = t1, obstacles interpolate along the frame.
bool make_collision_gv(const World& w, bool pt, const int gv[4], int obj_a,
                       int obj_b, double time, double dhat,
                       const std::vector<ObjInfo>& objs, bool force,
                       CollisionRec* out) {
  V3 x[4];
  double inv[4];
  for (int k = 0; k < 4; ++k) {
    x[k] = lerp(&w.pos_t0[3 * gv[k]], &w.pos_t1[3 * gv[k]], time);
    inv[k] = inv_mass_of(w, gv[k]);
  }
  CollisionRec r;
  if (!dcd_pair_forced(pt, gv, x, inv, dhat, force, &r)) return false;
  const float fa = objs[obj_a].friction, fb = objs[obj_b].friction;
  r.friction = std::sqrt(static_cast<double>(fa) * static_cast<double>(fb));
  // Prescribed obstacle displacement: coef_k (x_k(t1) - x_k(time)).
  for (int k = 0; k < 4; ++k) {
    if (gv[k] < w.nv_cloth) continue;
    for (int d = 0; d < 3; ++d) r.presc[d] += r.coef[k] * (w.pos_t1[3 * gv[k] + d] - x[k][d]);
  }
  *out = r;
  return true;
}

bool make_collision(const World& w, bool pt, const Prim* A, const Prim* B,
                    double time, double dhat, const std::vector<ObjInfo>& objs,
                    CollisionRec* out) {
  int gv[4];
  if (pt) {
    gv[0] = B->v[0];
    gv[1] = A->v[0];
    gv[2] = A->v[1];
    gv[3] = A->v[2];
  } else {
    gv[0] = A->v[0];
    gv[1] = A->v[1];
    gv[2] = B->v[0];
    gv[3] = B->v[1];
  }
  return make_collision_gv(w, pt, gv, A->obj, B->obj, time, dhat, objs, false, out);
}

// Object index (cloth c -> c, obstacle o -> ncloth + o) of a global vertex.
int object_of(const World& w, int gv) {
  const int c = cloth_of(w, gv);
  if (c >= 0) return c;
  for (size_t o = 0; o < w.obstacles.size(); ++o)
    if (gv >= w.obstacles[o].off && gv < w.obstacles[o].off + w.obstacles[o].nv)
      return static_cast<int>(w.cloths.size() + o);
  return -1;
}

std::array<int, 5> contact_key(const CollisionRec& r) {
  std::array<int, 5> k{r.type, r.gv[0], r.gv[1], r.gv[2], r.gv[3]};
  if (r.type == 0) {
    std::sort(k.begin() + 2, k.end());
  } else {
    if (k[2] < k[1]) std::swap(k[1], k[2]);
    if (k[4] < k[3]) std::swap(k[3], k[4]);
    if (std::make_pair(k[3], k[4]) < std::make_pair(k[1], k[2])) {
      std::swap(k[1], k[3]);
      std::swap(k[2], k[4]);
    }
  }
  return k;
}

}  // namespace

// ----------------------------------------------------------------- contacts

// Daviet contact of a DCD record (contact_constraints.cu:23-150): vertices
// with id >= nv_cloth are static (their terms go to affine); cbase_w holds
// the base contact weight per cloth dof.
Contact contact_of(const CollisionRec& r, int nv_cloth,
                   const std::vector<double>& cbase_w) {
  Contact c;
  c.mu = r.friction;
  double ref[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) {
    c.normal[d] = r.normal[d];
    c.affine[d] = -r.dhat * r.normal[d];
  }
  double scatter = 0.0;
  for (int k = 0; k < 4; ++k) {
    c.coef[k] = r.coef[k];
    for (int d = 0; d < 3; ++d) ref[d] += r.coef[k] * r.x[k][d];
    if (r.gv[k] >= nv_cloth) {
      c.state[k] = -1;
      for (int d = 0; d < 3; ++d) c.affine[d] += r.coef[k] * r.x[k][d];
      continue;
    }
    c.state[k] = r.gv[k];
    if (r.inv_mass[k] > 0.0) scatter += r.coef[k] * r.coef[k] / cbase_w[3 * r.gv[k]];
  }
  for (int d = 0; d < 3; ++d) c.affine[d] += r.presc[d];
  const double rn = ref[0] * c.normal[0] + ref[1] * c.normal[1] + ref[2] * c.normal[2];
  for (int d = 0; d < 3; ++d) c.affine[d] -= ref[d] - rn * c.normal[d];
  for (int k = 0; k < 4; ++k) {
    if (c.state[k] >= 0 && r.inv_mass[k] > 0.0)
      c.gamma[k] = c.coef[k] / (scatter * cbase_w[3 * c.state[k]]);
    else
      c.gamma[k] = 0.0;
  }
  return c;
}

// Greedy first-fit colouring in contact order over shared movable vertices;
// groups hold contact ids ascending per colour (contact_coloring.cu:241-392
// with a deterministic schedule).  A vertex takes part in a contact's solve
// when its coefficient is nonzero; in a replayed step the other solver's
// colours and vertex masks are taken instead (a barycentric coordinate
// clamped to exactly 0 on one side and not on the other changes both).
static void recolor(World& w) {
  const int n = static_cast<int>(w.contacts.size());
  const int nrp = w.replay ? static_cast<int>(w.rp_info.size()) : 0;
  std::vector<int> mask(n, 0);
  for (int c = 0; c < n; ++c) {
    Contact& ct = w.contacts[c];
    for (int k = 0; k < 4; ++k) {
      if (c < nrp) {
        const bool on = ((w.rp_info[c] >> k) & 1) != 0 && ct.state[k] >= 0;
        if (!on) ct.gamma[k] = 0.0;
        mask[c] |= on ? 1 << k : 0;
      } else {
        mask[c] |= ct.gamma[k] != 0.0 ? 1 << k : 0;
      }
    }
  }
  std::map<int, std::vector<int>> incid;
  for (int c = 0; c < n; ++c)
    for (int k = 0; k < 4; ++k)
      if ((mask[c] >> k) & 1) incid[w.contacts[c].state[k]].push_back(c);
  std::vector<int> color(n, 0);
  int maxc = 0;
  for (int c = 0; c < n; ++c) {
    int col = 1;
    if (c < nrp) {
      col = std::max(1, w.rp_info[c] >> 4);
    } else {
      std::set<int> used;
      for (int k = 0; k < 4; ++k) {
        if (((mask[c] >> k) & 1) == 0) continue;
        for (int o : incid[w.contacts[c].state[k]])
          if (o < c) used.insert(color[o]);
      }
      while (used.count(col)) ++col;
    }
    color[c] = col;
    maxc = std::max(maxc, col);
  }
  w.log_info.resize(n);
  for (int c = 0; c < n; ++c) w.log_info[c] = color[c] << 4 | mask[c];
  w.color_groups.assign(maxc, {});
  for (int c = 0; c < n; ++c) w.color_groups[color[c] - 1].push_back(c);
  // active weights
  std::fill(w.cw.begin(), w.cw.end(), 0.0);
  int active_v = 0;
  for (auto& [v, list] : incid) {
    (void)list;
    ++active_v;
    for (int d = 0; d < 3; ++d) w.cw[3 * v + d] = w.cbase_w[3 * v + d];
  }
  w.active_dof = 3 * active_v;
}

static int append_collisions(World& w, std::vector<CollisionRec> recs, int call) {
  std::sort(recs.begin(), recs.end(), [](const CollisionRec& a, const CollisionRec& b) {
    return contact_key(a) < contact_key(b);
  });
  int added = 0;
  for (const auto& r : recs) {
    if (!w.keys.insert(contact_key(r)).second) continue;
    w.log_contacts.insert(w.log_contacts.end(),
                          {call, r.type, r.gv[0], r.gv[1], r.gv[2], r.gv[3]});
    w.contacts.push_back(contact_of(r, w.nv_cloth, w.cbase_w));
    ++added;
  }
  if (added > 0) recolor(w);
  return added;
}

static std::vector<CollisionRec> dcd(World& w, double time, double dhat) {
  auto objs = build_objects(w);
  auto cand = candidates(w, objs);
  std::vector<CollisionRec> out;
  CollisionRec r;
  for (auto& [t, p] : cand.pt)
    if (make_collision(w, true, t, p, time, dhat, objs, &r)) out.push_back(r);
  for (auto& [a, b] : cand.ee)
    if (make_collision(w, false, a, b, time, dhat, objs, &r)) out.push_back(r);
  return out;
}

// Replay: the records of the contacts another solver appended at DCD call
// `call`, formed on this world's positions whatever its own activation test
// says (a projection just outside the triangle or a distance just above
// dhat is kept).
static std::vector<CollisionRec> dcd_replay(World& w, int call, double time,
                                            double dhat) {
  auto objs = build_objects(w);
  std::vector<CollisionRec> out;
  CollisionRec r;
  for (size_t i = 0; i + 6 <= w.rp_contacts.size(); i += 6) {
    const int* e = &w.rp_contacts[i];
    if (e[0] != call) continue;
    const bool pt = e[1] == 0;
    const int gv[4] = {e[2], e[3], e[4], e[5]};
    bool valid = true;
    for (int k = 0; k < 4; ++k) valid = valid && gv[k] >= 0 && gv[k] < w.nv_total;
    const int oa = valid ? object_of(w, pt ? gv[1] : gv[0]) : -1;
    const int ob = valid ? object_of(w, pt ? gv[0] : gv[2]) : -1;
    if (oa < 0 || ob < 0 ||
        !make_collision_gv(w, pt, gv, oa, ob, time, dhat, objs, true, &r)) {
      ++w.replay_missing;
      continue;
    }
    out.push_back(r);
  }
  return out;
}

static double min_toi(World& w, double time_start, double interval, double ms) {
  auto objs = build_objects(w);
  auto cand = candidates(w, objs);
  w.stats.ccd_cand += static_cast<int>(cand.pt.size() + cand.ee.size());
  double abs_max[3] = {0, 0, 0};
  for (int i = 0; i < w.nv_total; ++i)
    for (int d = 0; d < 3; ++d)
      abs_max[d] = std::max({abs_max[d],
                             std::fabs(static_cast<double>(static_cast<float>(w.pos_t0[3 * i + d]))),
                             std::fabs(static_cast<double>(static_cast<float>(w.pos_t1[3 * i + d])))});
  CcdConfig cfg;
  cfg.ms = ms;
  cfg.tolerance = w.P.ccd_tol / interval;
  cfg.max_iter = w.P.ccd_max_iter;
  cfg.max_time = 1.0;
  cfg.stack_cap = kCcdStackCap;
  cfg.zero_mode = w.P.ccd_zero;
  double best = 1.0;
  auto traj = [&](int gv, double p0[3], double p1[3]) {
    if (gv < w.nv_cloth) {
      for (int d = 0; d < 3; ++d) {
        p0[d] = w.pos_t0[3 * gv + d];
        p1[d] = w.pos_t1[3 * gv + d];
      }
    } else {
      for (int d = 0; d < 3; ++d) {
        const double a = w.pos_t0[3 * gv + d], b = w.pos_t1[3 * gv + d];
        const double delta = b - a;
        p0[d] = a + time_start * delta;
        p1[d] = p0[d] + interval * delta;
      }
    }
  };
  auto run = [&](bool vf, const int gv[4]) {
    double p0[4][3], p1[4][3];
    for (int k = 0; k < 4; ++k) traj(gv[k], p0[k], p1[k]);
    CcdConfig c = cfg;
    ccd_filter(abs_max, vf, c.err);
    double toi;
    bool zero = false;
    if (!ccd_query(vf, p0, p1, c, &toi, &zero)) {
      w.stats.ccd_zero += zero ? 1 : 0;
      return;
    }
    if (toi == 0.0) return;
    best = std::min(best, toi);
  };
  for (auto& [t, p] : cand.pt) {
    int gv[4] = {p->v[0], t->v[0], t->v[1], t->v[2]};
    run(true, gv);
  }
  for (auto& [a, b] : cand.ee) {
    int gv[4] = {a->v[0], a->v[1], b->v[0], b->v[1]};
    run(false, gv);
  }
  return interval * best;
}

// Contact ADMM update (contact_constraints.cu:203-264,391-411).
static void contact_update(World& w, const std::vector<double>& x,
                           std::vector<double>& dual, std::vector<double>& dcur,
                           std::vector<double>& dprev, double* primal) {
  if (w.contacts.empty()) return;
  w.p_prev = w.p;
  for (size_t i = 0; i < x.size(); ++i)
    w.p[i] = w.cw[i] > 0.0 ? x[i] - w.eta[i] / w.cw[i] : x[i];
  for (auto& grp : w.color_groups)
    for (int ci : grp) {
      Contact& c = w.contacts[ci];
      for (int k = 0; k < 4; ++k)
        if (c.gamma[k] != 0.0)
          for (int d = 0; d < 3; ++d) w.p[3 * c.state[k] + d] += c.gamma[k] * c.corr[d];
    }
  for (int s = 0; s < w.P.sweeps; ++s)
    for (auto& grp : w.color_groups)
      for (int ci : grp) {
        Contact& c = w.contacts[ci];
        double uu[3] = {c.affine[0], c.affine[1], c.affine[2]};
        for (int k = 0; k < 4; ++k)
          if (c.state[k] >= 0)
            for (int d = 0; d < 3; ++d) uu[d] += c.coef[k] * w.p[3 * c.state[k] + d];
        double us[3];
        for (int d = 0; d < 3; ++d) us[d] = uu[d] - c.corr[d];
        double unew[3];
        coulomb_solve(us, c.normal, c.mu, unew);
        double du[3];
        for (int d = 0; d < 3; ++d) {
          du[d] = unew[d] - uu[d];
          c.corr[d] += du[d];
        }
        for (int k = 0; k < 4; ++k)
          if (c.gamma[k] != 0.0)
            for (int d = 0; d < 3; ++d) w.p[3 * c.state[k] + d] += c.gamma[k] * du[d];
      }
  for (size_t i = 0; i < x.size(); ++i) {
    if (!(w.cw[i] > 0.0)) continue;
    const double r = x[i] - w.p[i];
    w.eta[i] += w.cw[i] * (w.p[i] - x[i]);
    *primal += r * r;
    dual[i] += w.cw[i] * (w.p[i] - w.p_prev[i]);
    dcur[i] += w.eta[i];
    dprev[i] += w.eta[i];
  }
}

// ----------------------------------------------------------------- ADMM

// `forced` >= 0 runs exactly that many iterations (a replayed decision of
// the convergence test below); -1 stops on the test.
static bool admm_solve(World& w, std::vector<double>& X,
                       const std::vector<double>& minert, int forced = -1) {
  const int N = 3 * w.nv_cloth;
  std::vector<double> lam(N, 0.0);
  double init_p = 0, init_d = 0, ratio = 1.0;
  int faces = 0, pin_dof = 0;
  for (auto& c : w.cloths) {
    faces += c.nf;
    for (char p : c.pinned) pin_dof += p ? 3 : 0;
  }
  const double rho_p = 1e4, max_mul = 1e8;
  const int n_it = forced >= 0 ? forced : w.P.max_inner;
  for (int it = 0; it < n_it; ++it) {
    if (it != 0) {
      const double rel = std::clamp(w.P.lin_rel_init * ratio, w.P.lin_rel_min,
                                    w.P.lin_rel_max);
      for (auto& c : w.cloths) {
        const int o = c.off;
        std::vector<double> diag(3 * c.nv), b(3 * c.nv), xs(3 * c.nv);
        for (int i = 0; i < c.nv; ++i)
          for (int d = 0; d < 3; ++d) {
            const int gi = 3 * (o + i) + d, li = 3 * i + d;
            double dd = c.mass[i] / (w.P.dt * w.P.dt);
            double bb = c.bend[li] + minert[gi];
            if (c.pinned[i]) {
              dd += rho_p;
              bb += rho_p * c.pin_target[li] - lam[gi];
            }
            if (w.cw[gi] > 0.0) {
              dd += w.cw[gi];
              bb += w.cw[gi] * w.p[gi] + w.eta[gi];
            }
            diag[li] = dd;
            b[li] = bb;
            xs[li] = X[gi];
          }
        // elastic rhs: sum_f S^T (rho w^2 z - w u)
        for (int f = 0; f < c.nf; ++f) {
          const double wgt = std::sqrt(c.l2.area[f]);
          const double rho = c.k;
          const auto& J = c.l2.jop[f];
          for (int col = 0; col < 9; ++col) {
            double s = 0.0;
            for (int r = 0; r < 6; ++r)
              s += J[9 * r + col] * (rho * wgt * wgt * c.z[6 * f + r] - wgt * c.u[6 * f + r]);
            b[3 * c.F[3 * f + col / 3] + col % 3] += s;
          }
        }
        if (it == 1) c.mas.factorize(c.A, diag);
        auto res = pcg_solve(c.A, diag, c.mas, b.data(), xs.data(), rel,
                             w.P.lin_abs, w.P.pcg_max_iter, w.P.pcg_criterion);
        w.stats.pcg += res.iters;
        for (int i = 0; i < 3 * c.nv; ++i) X[3 * o + i] = xs[i];
      }
    }
    // aux update
    double s_primal = 0, s_x2 = 0, s_z2 = 0;
    std::vector<double> dual(N, 0.0), dcur(N, 0.0), dprev(N, 0.0);
    for (auto& c : w.cloths) {
      for (int f = 0; f < c.nf; ++f) {
        const auto& J = c.l2.jop[f];
        double xl[9];
        for (int k = 0; k < 3; ++k)
          for (int d = 0; d < 3; ++d) xl[3 * k + d] = X[3 * (c.off + c.F[3 * f + k]) + d];
        double Sx[6], y[6], z[6], proj[6];
        const double wgt = std::sqrt(c.l2.area[f]), rho = c.k, k = c.k;
        for (int r = 0; r < 6; ++r) {
          Sx[r] = 0.0;
          for (int col = 0; col < 9; ++col) Sx[r] += J[9 * r + col] * xl[col];
          y[r] = Sx[r] + c.u[6 * f + r] / (rho * wgt);
        }
        project_stretch(y, 0.9, 1.1, proj);
        const double rr = k / (k + rho);
        for (int r = 0; r < 6; ++r) z[r] = (1.0 - rr) * y[r] + rr * proj[r];
        for (int r = 0; r < 6; ++r) {
          const double pr = wgt * (Sx[r] - z[r]);
          s_primal += pr * pr;
          s_x2 += (wgt * Sx[r]) * (wgt * Sx[r]);
          s_z2 += (wgt * z[r]) * (wgt * z[r]);
        }
        for (int col = 0; col < 9; ++col) {
          double a = 0.0, b = 0.0;
          for (int r = 0; r < 6; ++r) {
            a += J[9 * r + col] * z[r];
            b += J[9 * r + col] * c.z[6 * f + r];
          }
          const int gi = 3 * (c.off + c.F[3 * f + col / 3]) + col % 3;
          const double s = rho * wgt * wgt;
          dual[gi] += s * a - s * b;
          dcur[gi] += s * a;
          dprev[gi] += s * b;
        }
        for (int r = 0; r < 6; ++r) {
          c.z[6 * f + r] = z[r];
          double nu = c.u[6 * f + r] + wgt * rho * (Sx[r] - z[r]);
          c.u[6 * f + r] = std::clamp(nu, -max_mul, max_mul);
        }
      }
      for (int i = 0; i < c.nv; ++i) {
        if (!c.pinned[i]) continue;
        for (int d = 0; d < 3; ++d) {
          const int gi = 3 * (c.off + i) + d;
          const double t = c.pin_target[3 * i + d];
          lam[gi] = std::clamp(lam[gi] + rho_p * (X[gi] - t), -max_mul, max_mul);
          s_primal += (X[gi] - t) * (X[gi] - t);
          s_x2 += X[gi] * X[gi];
          s_z2 += t * t;
        }
      }
    }
    contact_update(w, X, dual, dcur, dprev, &s_primal);
    w.stats.admm++;
    if (it == 0) continue;
    for (double v : X)
      if (!std::isfinite(v)) return false;
    const int dof_p = 6 * faces + pin_dof + w.active_dof;
    double dn = 0, dc = 0, dp = 0;
    for (int i = 0; i < N; ++i) {
      dn += dual[i] * dual[i];
      dc += dcur[i] * dcur[i];
      dp += dprev[i] * dprev[i];
    }
    const double primal = std::sqrt(s_primal), dualn = std::sqrt(dn);
    if (it == 1) {
      init_p = primal;
      init_d = dualn;
    }
    const double eps_p = std::sqrt(double(dof_p)) * w.P.admm_abs +
                         std::max(std::sqrt(s_x2), std::sqrt(s_z2)) * w.P.admm_rel;
    const double eps_d = std::sqrt(double(N)) * w.P.admm_abs +
                         std::max(std::sqrt(dc), std::sqrt(dp)) * w.P.admm_rel;
    ratio = std::max(primal / std::max(init_p, w.P.admm_abs),
                     dualn / std::max(init_d, w.P.admm_abs));
    if (forced < 0 && primal <= eps_p && dualn <= eps_d) break;
  }
  for (auto& c : w.cloths)
    for (int i = 0; i < c.nv; ++i)
      if (c.pinned[i])
        for (int d = 0; d < 3; ++d) X[3 * (c.off + i) + d] = c.pin_target[3 * i + d];
  for (int i = 0; i < N; ++i)
    if (w.cw[i] > 0.0) X[i] = w.p[i];
  return true;
}

static bool step(World& w) {
  w.stats = StepStats{};
  if (w.cloths.empty()) return false;
  const double h = w.P.dt;
  // offsets
  int off = 0;
  for (auto& c : w.cloths) {
    c.off = off;
    off += c.nv;
  }
  w.nv_cloth = off;
  for (auto& o : w.obstacles) {
    o.off = off;
    off += o.nv;
  }
  w.nv_total = off;
  const int N = 3 * w.nv_cloth;
  std::vector<double> prev_x(N), prev_v(N), minert(N);
  for (auto& c : w.cloths)
    for (int i = 0; i < c.nv; ++i)
      for (int d = 0; d < 3; ++d) {
        const int gi = 3 * (c.off + i) + d;
        prev_x[gi] = c.x[3 * i + d];
        prev_v[gi] = (1.0 - c.damping) * c.v[3 * i + d];
        const double mod = -prev_x[gi] / (h * h) - prev_v[gi] / h - w.P.acc[d];
        minert[gi] = -c.mass[i] * mod;
      }
  for (auto& c : w.cloths) std::fill(c.z.begin(), c.z.end(), 0.0), std::fill(c.u.begin(), c.u.end(), 0.0);
  std::vector<double> outer_x = prev_x, outer_v = prev_v, inner(N);
  // positions of all vertices
  w.pos_t0.assign(3 * w.nv_total, 0.0);
  w.pos_t1.assign(3 * w.nv_total, 0.0);
  for (auto& o : w.obstacles)
    for (int i = 0; i < 3 * o.nv; ++i) {
      w.pos_t0[3 * o.off + i] = o.prev[i];
      w.pos_t1[3 * o.off + i] = o.curr[i];
    }
  auto set_cloth = [&](const std::vector<double>& a, const std::vector<double>& b) {
    for (int i = 0; i < N; ++i) {
      w.pos_t0[i] = a[i];
      w.pos_t1[i] = b[i];
    }
  };
  float pad = 1e30f;
  for (auto& c : w.cloths) pad = std::min(pad, c.padding);
  for (auto& o : w.obstacles) pad = std::min(pad, o.padding);
  const double ms = w.P.ms_scale * pad, dhat = w.P.dcd_scale * pad;
  // contacts
  w.contacts.clear();
  w.keys.clear();
  w.color_groups.clear();
  w.cbase_w.assign(N, 0.0);
  for (auto& c : w.cloths)
    for (int i = 0; i < 3 * c.nv; ++i) w.cbase_w[3 * c.off + i] = c.cw[i];
  w.cw.assign(N, 0.0);
  w.p.assign(N, 0.0);
  w.p_prev.assign(N, 0.0);
  w.eta.assign(N, 0.0);
  w.active_dof = 0;
  w.log_toi.clear();
  w.log_admm.clear();
  w.log_info.clear();
  w.log_contacts.clear();
  w.replay_missing = 0;
  int call = 0;
  auto find_contacts = [&](double time) {
    append_collisions(w, w.replay ? dcd_replay(w, call, time, dhat) : dcd(w, time, dhat),
                      call);
    ++call;
  };
  set_cloth(outer_x, outer_x);
  find_contacts(0.0);
  double remaining = 1.0;
  // A replayed step also replays the other solver's acceptance test, which
  // that solver makes in FP32 (toi >= remaining, remaining -= 0.8 toi): in
  // FP64 the final TOI, equal to that solver's FP32 remaining interval, can
  // fall just below this side's remaining interval and start an outer
  // iteration the other solver never ran.
  float remaining_f = 1.0f;
  int outer = 0;
  for (; outer < w.P.max_outer; ++outer) {
    for (int i = 0; i < N; ++i) {
      const int d = i % 3;
      inner[i] = (outer_x[i] + h * outer_v[i]) + h * h * w.P.acc[d];
    }
    const int forced = w.replay && outer < static_cast<int>(w.rp_admm.size())
                           ? w.rp_admm[outer]
                           : -1;
    const int admm0 = w.stats.admm;
    if (!admm_solve(w, inner, minert, forced)) return false;
    w.log_admm.push_back(static_cast<int>(w.stats.admm - admm0));
    set_cloth(outer_x, inner);
    const bool replayed = w.replay && outer < static_cast<int>(w.rp_toi.size());
    const double toi =
        replayed ? w.rp_toi[outer] : min_toi(w, 1.0 - remaining, remaining, ms);
    w.log_toi.push_back(toi);
    for (int i = 0; i < N; ++i) outer_v[i] = (inner[i] - outer_x[i]) / h;
    const bool accept =
        replayed ? static_cast<float>(toi) >= remaining_f : toi >= remaining;
    if (accept) {
      outer_x = inner;
      break;
    }
    remaining_f -= 0.8f * static_cast<float>(toi);
    w.stats.min_toi = std::min(w.stats.min_toi, 1.0 - remaining + toi);
    const double rb = 0.8 * toi, frac = rb / remaining;
    for (int i = 0; i < N; ++i) outer_x[i] = frac * inner[i] + (1.0 - frac) * outer_x[i];
    remaining -= rb;
    set_cloth(outer_x, outer_x);
    find_contacts(1.0 - remaining);
  }
  w.replay = false;
  w.stats.outer = std::min(outer + 1, w.P.max_outer);
  w.stats.contacts = static_cast<int>(w.contacts.size());
  w.stats.colors = static_cast<int>(w.color_groups.size());
  for (auto& c : w.cloths)
    for (int i = 0; i < 3 * c.nv; ++i) {
      const int gi = 3 * c.off + i;
      c.v[i] = (outer_x[gi] - prev_x[gi]) / h;
      c.x[i] = outer_x[gi];
    }
  for (auto& o : w.obstacles) {
    o.prev = o.curr;
    if (o.is_static) o.is_static_twice = true;
    else o.is_static = true;
  }
  return true;
}

}  // namespace oracle

// ------------------------------------------------------------------- C API

using oracle::World;

extern "C" {

void* oracle_create() { return new World(); }
void oracle_destroy(void* w) { delete static_cast<World*>(w); }

/// p: 20 doubles: acc[3], max_outer, max_inner, dt, lin_abs, lin_rel_min,
/// lin_rel_max, lin_rel_init, admm_abs, admm_rel, pcg_max_iter, sweeps,
/// ccd_max_iter, ccd_tol, ms_scale, dcd_scale, pcg_criterion, ccd_zero.
void oracle_set_global(void* wp, const double* p) {
  auto& P = static_cast<World*>(wp)->P;
  P.acc[0] = p[0]; P.acc[1] = p[1]; P.acc[2] = p[2];
  P.max_outer = static_cast<int>(p[3]);
  P.max_inner = static_cast<int>(p[4]);
  P.dt = p[5];
  P.lin_abs = p[6]; P.lin_rel_min = p[7]; P.lin_rel_max = p[8]; P.lin_rel_init = p[9];
  P.admm_abs = p[10]; P.admm_rel = p[11];
  P.pcg_max_iter = static_cast<int>(p[12]);
  P.sweeps = static_cast<int>(p[13]);
  P.ccd_max_iter = static_cast<int>(p[14]);
  P.ccd_tol = p[15]; P.ms_scale = p[16]; P.dcd_scale = p[17];
  P.pcg_criterion = static_cast<int>(p[18]);
  P.ccd_zero = static_cast<int>(p[19]);
}

/// cloth: k, kb, density, damping; coll: on, self, group, restitution,
/// friction.  perm / offsets: the solver partition (caller -> solver order).
int oracle_add_cloth(void* wp, const double* cloth, const double* coll,
                     const float* V, int nv, const int* F, int nf, const int* pins,
                     int npin, const int* perm, const int* offsets, int nparts) {
  World& w = *static_cast<World*>(wp);
  oracle::OCloth c;
  c.k = cloth[0]; c.kb = cloth[1]; c.density = cloth[2]; c.damping = cloth[3];
  c.coll.on = static_cast<int>(coll[0]); c.coll.self_on = static_cast<int>(coll[1]);
  c.coll.group = static_cast<int>(coll[2]);
  c.coll.restitution = static_cast<float>(coll[3]);
  c.coll.friction = static_cast<float>(coll[4]);
  c.nv = nv;
  c.nf = nf;
  c.perm.assign(perm, perm + nv);
  c.part_offsets.assign(offsets, offsets + nparts + 1);
  std::vector<float> Vc(V, V + 3 * nv);
  std::vector<int> Fc(F, F + 3 * nf), Ec;
  float avg;
  oracle::edges_and_avg(Vc, Fc, Ec, avg);
  c.padding = 0.05f * avg;
  c.V0.assign(3 * nv, 0.0);
  for (int i = 0; i < nv; ++i)
    for (int d = 0; d < 3; ++d) c.V0[3 * perm[i] + d] = V[3 * i + d];
  c.F.resize(3 * nf);
  for (int i = 0; i < 3 * nf; ++i) c.F[i] = perm[F[i]];
  c.E.resize(Ec.size());
  for (size_t i = 0; i < Ec.size(); ++i) c.E[i] = perm[Ec[i]];
  c.pinned.assign(nv, 0);
  c.pin_target.assign(3 * nv, 0.0);
  for (int i = 0; i < npin; ++i) {
    const int lv = perm[pins[i]];
    c.pinned[lv] = 1;
    for (int d = 0; d < 3; ++d) c.pin_target[3 * lv + d] = V[3 * pins[i] + d];
  }
  c.l2 = oracle::build_l2(c.V0, c.F);
  c.A = oracle::build_operator(c.l2, c.F, c.k, c.kb);
  c.mas.build_topology(c.A, c.part_offsets);
  c.mass.resize(nv);
  for (int i = 0; i < nv; ++i) c.mass[i] = c.density * c.l2.mass[i];
  c.bend.resize(3 * nv);
  for (int i = 0; i < 3 * nv; ++i) c.bend[i] = c.kb * c.l2.c0[i];
  c.x = c.V0;
  c.v.assign(3 * nv, 0.0);
  c.z.assign(6 * nf, 0.0);
  c.u.assign(6 * nf, 0.0);
  w.cloths.push_back(std::move(c));
  return static_cast<int>(w.cloths.size()) - 1;
}

/// Contact weights depend on dt: call after oracle_set_global.
void oracle_finalize(void* wp) {
  World& w = *static_cast<World*>(wp);
  const double h = w.P.dt;
  for (auto& c : w.cloths) {
    c.cw.assign(3 * c.nv, 0.0);
    for (int i = 0; i < c.nv; ++i) {
      double blk[9] = {};
      auto it = c.A.rows[i].find(i);
      if (it != c.A.rows[i].end())
        for (int e = 0; e < 9; ++e) blk[e] = it->second[e];
      const double inertia = c.mass[i] / (h * h);
      for (int d = 0; d < 3; ++d) blk[4 * d] += inertia;
      const double eta = oracle::sym3_min_eig(blk);
      const double wt = std::max(0.001 * eta, std::min(25.0 * inertia, eta));
      for (int d = 0; d < 3; ++d) c.cw[3 * i + d] = wt;
    }
  }
}

int oracle_add_obstacle(void* wp, const double* coll, const float* V, int nv,
                        const int* F, int nf) {
  World& w = *static_cast<World*>(wp);
  oracle::OObstacle o;
  o.coll.on = static_cast<int>(coll[0]); o.coll.self_on = static_cast<int>(coll[1]);
  o.coll.group = static_cast<int>(coll[2]);
  o.coll.restitution = static_cast<float>(coll[3]);
  o.coll.friction = static_cast<float>(coll[4]);
  o.nv = nv;
  o.curr.assign(V, V + 3 * nv);
  o.prev = o.curr;
  o.F.assign(F, F + 3 * nf);
  float avg;
  oracle::edges_and_avg(o.curr, o.F, o.E, avg);
  o.padding = 0.05f * avg;
  w.obstacles.push_back(std::move(o));
  return static_cast<int>(w.obstacles.size()) - 1;
}

void oracle_set_obstacle_position(void* wp, int idx, const float* V) {
  World& w = *static_cast<World*>(wp);
  auto& o = w.obstacles[idx];
  if (!o.has_pos) {
    // First position record: static at that position
    // (cuda_backend.cu:339-349, PinPosition::from_position).
    o.curr.assign(V, V + 3 * o.nv);
    o.prev = o.curr;
    o.is_static = true;
    o.is_static_twice = true;
    o.has_pos = true;
    return;
  }
  o.is_static = false;
  o.is_static_twice = false;
  o.prev = o.curr;
  o.curr.assign(V, V + 3 * o.nv);
}

/// Obstacle idx moves from prev to curr over the next step (as after two
/// position records; both 3 * nv floats).
void oracle_set_obstacle_frame(void* wp, int idx, const float* prev, const float* curr) {
  World& w = *static_cast<World*>(wp);
  auto& o = w.obstacles[idx];
  o.prev.assign(prev, prev + 3 * o.nv);
  o.curr.assign(curr, curr + 3 * o.nv);
  o.has_pos = true;
  o.is_static = false;
  o.is_static_twice = false;
}

/// Replay for the next step: the outer loop takes tois[k] as the CCD result
/// of iteration k (frame-time units) and the DCD call c appends exactly the
/// contacts listed with call index c (6 ints each: call, type, four global
/// vertex ids in record order).
void oracle_set_replay(void* wp, const double* tois, const int* admm, int n_toi,
                       const int* contacts, const int* info, int n_contacts) {
  World& w = *static_cast<World*>(wp);
  w.replay = true;
  w.rp_toi.assign(tois, tois + n_toi);
  if (admm != nullptr) w.rp_admm.assign(admm, admm + n_toi);
  else w.rp_admm.clear();
  if (info != nullptr) w.rp_info.assign(info, info + n_contacts);
  else w.rp_info.clear();
  w.rp_contacts.assign(contacts, contacts + 6 * static_cast<size_t>(n_contacts));
}

/// Decisions of the last step in the replay format; counts in n_toi and
/// n_contacts (buffers may be null).  Returns the number of replayed
/// contacts whose record could not be formed.
int oracle_decisions(void* wp, double* tois, int* admm, int cap_toi, int* n_toi,
                     int* contacts, int* info, int cap_contacts, int* n_contacts) {
  World& w = *static_cast<World*>(wp);
  *n_toi = static_cast<int>(w.log_toi.size());
  *n_contacts = static_cast<int>(w.log_contacts.size() / 6);
  if (tois)
    for (int i = 0; i < std::min(cap_toi, *n_toi); ++i) tois[i] = w.log_toi[i];
  if (admm)
    for (int i = 0; i < std::min(cap_toi, *n_toi); ++i) admm[i] = w.log_admm[i];
  if (contacts)
    for (size_t i = 0; i < 6 * static_cast<size_t>(std::min(cap_contacts, *n_contacts)); ++i)
      contacts[i] = w.log_contacts[i];
  if (info)
    for (int i = 0; i < std::min(cap_contacts, *n_contacts); ++i) info[i] = w.log_info[i];
  return w.replay_missing;
}

/// DCD records of the current objects on explicit positions (global
/// numbering, 3 per vertex per end) at `time`: 6 ints per record (type, four
/// global vertex ids in record order, 0) and its distance.  Returns 0, or 1
/// if cap was exceeded (the count is still exact).
int oracle_dcd(void* wp, const double* pos0, const double* pos1, double time,
               double dhat, int* recs, double* dist, int cap, int* n) {
  World& w = *static_cast<World*>(wp);
  int off = 0;
  for (auto& c : w.cloths) {
    c.off = off;
    off += c.nv;
  }
  w.nv_cloth = off;
  for (auto& o : w.obstacles) {
    o.off = off;
    off += o.nv;
  }
  w.nv_total = off;
  w.pos_t0.assign(pos0, pos0 + 3 * static_cast<size_t>(off));
  w.pos_t1.assign(pos1, pos1 + 3 * static_cast<size_t>(off));
  const std::vector<oracle::CollisionRec> out = oracle::dcd(w, time, dhat);
  *n = static_cast<int>(out.size());
  for (int i = 0; i < std::min(cap, *n); ++i) {
    const auto& r = out[i];
    int* e = recs + 6 * static_cast<size_t>(i);
    e[0] = r.type;
    for (int k = 0; k < 4; ++k) e[1 + k] = r.gv[k];
    e[5] = 0;
    double d = 0.0;
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) d += r.coef[k] * r.x[k][c] * r.normal[c];
    dist[i] = d;
  }
  return *n > cap ? 1 : 0;
}

/// stats: outer, admm, pcg, contacts, colors, ccd candidates, min_toi,
/// candidates left with a root at t = 0.
int oracle_step(void* wp, double* stats) {
  World& w = *static_cast<World*>(wp);
  const bool ok = oracle::step(w);
  if (stats) {
    stats[0] = w.stats.outer; stats[1] = w.stats.admm; stats[2] = w.stats.pcg;
    stats[3] = w.stats.contacts; stats[4] = w.stats.colors;
    stats[5] = w.stats.ccd_cand; stats[6] = w.stats.min_toi;
    stats[7] = w.stats.ccd_zero;
  }
  return ok ? 0 : 1;
}

/// Broadphase candidates of the current objects for explicit segment end
/// positions pos0 -> pos1 (global numbering: cloths in solver order, then
/// obstacles).  PT entries are (point, tri0, tri1, tri2), EE entries
/// (a0, a1, b0, b1) in canonical edge order.  Returns 0, or 1 if a list
/// exceeded its capacity (counts are still exact).
int oracle_candidates(void* wp, const double* pos0, const double* pos1,
                      int* pt, int pt_cap, int* npt, int* ee, int ee_cap,
                      int* nee) {
  World& w = *static_cast<World*>(wp);
  int off = 0;
  for (auto& c : w.cloths) {
    c.off = off;
    off += c.nv;
  }
  w.nv_cloth = off;
  for (auto& o : w.obstacles) {
    o.off = off;
    off += o.nv;
  }
  w.nv_total = off;
  w.pos_t0.assign(pos0, pos0 + 3 * off);
  w.pos_t1.assign(pos1, pos1 + 3 * off);
  auto objs = oracle::build_objects(w);
  auto cand = oracle::candidates(w, objs);
  *npt = static_cast<int>(cand.pt.size());
  *nee = static_cast<int>(cand.ee.size());
  int i = 0;
  for (auto& [t, p] : cand.pt) {
    if (i < pt_cap) {
      pt[4 * i] = p->v[0];
      for (int k = 0; k < 3; ++k) pt[4 * i + 1 + k] = t->v[k];
    }
    ++i;
  }
  i = 0;
  for (auto& [a, b] : cand.ee) {
    if (i < ee_cap) {
      ee[4 * i] = a->v[0];
      ee[4 * i + 1] = a->v[1];
      ee[4 * i + 2] = b->v[0];
      ee[4 * i + 3] = b->v[1];
    }
    ++i;
  }
  return (*npt > pt_cap || *nee > ee_cap) ? 1 : 0;
}

/// Positions of cloth `idx` in caller order (double).
void oracle_get_positions(void* wp, int idx, double* out) {
  World& w = *static_cast<World*>(wp);
  const auto& c = w.cloths[idx];
  for (int i = 0; i < c.nv; ++i)
    for (int d = 0; d < 3; ++d) out[3 * i + d] = c.x[3 * c.perm[i] + d];
}

/// The DCD record and Daviet contact of one pair on explicit positions
/// (x12: 4 vertices; PT: point, tri0..2; EE: a0, a1, b0, b1).  Vertices
/// with is_static[k] have no state (their terms go to affine); inv_mass[k] = 0
/// marks a pinned cloth vertex; cw4: base contact weight of each vertex.
/// out (37): dist, u-param, v-param, normal[3], coef[4], gamma[4],
/// state_valid[4], affine[3], and u = sum coef x + affine at x12 [3], then
/// zeros.  Returns 0 when the pair makes no contact.
int oracle_dcd_contact(int pt, const double* x12, const double* inv_mass4,
                       const int* is_static4, const double* cw4, double dhat,
                       double* out) {
  using namespace oracle;
  int gv[4];
  V3 x[4];
  double inv[4];
  std::vector<double> cbase(3 * 8, 0.0);
  for (int k = 0; k < 4; ++k) {
    gv[k] = is_static4[k] ? 4 + k : k;
    for (int d = 0; d < 3; ++d) x[k][d] = x12[3 * k + d];
    inv[k] = is_static4[k] ? 0.0 : inv_mass4[k];
    for (int d = 0; d < 3; ++d) cbase[3 * gv[k] + d] = cw4[k];
  }
  CollisionRec r;
  if (!dcd_pair(pt != 0, gv, x, inv, dhat, &r)) return 0;
  r.friction = 0.3;
  Contact c = contact_of(r, 4, cbase);
  std::fill(out, out + 37, 0.0);
  V3 disp{0, 0, 0};
  for (int k = 0; k < 4; ++k)
    for (int d = 0; d < 3; ++d) disp[d] += r.coef[k] * x[k][d];
  out[0] = std::sqrt(dot(disp, disp));
  out[1] = pt ? -r.coef[2] : r.coef[1];
  out[2] = pt ? -r.coef[3] : -r.coef[3];
  for (int d = 0; d < 3; ++d) out[3 + d] = r.normal[d];
  for (int k = 0; k < 4; ++k) {
    out[6 + k] = c.coef[k];
    out[10 + k] = c.gamma[k];
    out[14 + k] = c.state[k] >= 0 ? 1.0 : 0.0;
  }
  for (int d = 0; d < 3; ++d) out[18 + d] = c.affine[d];
  for (int d = 0; d < 3; ++d) {
    double u = c.affine[d];
    for (int k = 0; k < 4; ++k)
      if (c.state[k] >= 0) u += c.coef[k] * x[k][d];
    out[21 + d] = u;
  }
  return 1;
}

/// Smallest eigenvalue of a symmetric 3x3 matrix (the contact weight's eta,
/// cloth_assembly_l1_cache.cu:60-78).
double oracle_sym3_min_eig(const double* a9) { return oracle::sym3_min_eig(a9); }

/// The contact solve of one ADMM iteration on explicit contacts
/// (contact_constraints.cu:203-411 as the oracle step runs it): nc contacts
/// over nv cloth vertices, state (4 per contact, -1 = static), coef, gamma
/// (4), normal, affine (3), mu (1); x (3nv), base weight cw (nv; a vertex is
/// active when some contact moves it), eta (3nv, in/out), corr (3nc,
/// in/out).  Writes p (3nv) and the greedy colour of each contact (from 1).
void oracle_contact_sweeps(int nc, const int* state, const double* coef,
                           const double* gamma, const double* normal,
                           const double* affine, const double* mu, int nv,
                           const double* x, const double* cw, double* eta,
                           double* corr, int sweeps, double* p_out,
                           int* color_out) {
  using namespace oracle;
  World w;
  w.P.sweeps = sweeps;
  w.nv_cloth = nv;
  const int N = 3 * nv;
  w.cbase_w.assign(N, 0.0);
  for (int i = 0; i < nv; ++i)
    for (int d = 0; d < 3; ++d) w.cbase_w[3 * i + d] = cw[i];
  w.cw.assign(N, 0.0);
  w.p.assign(N, 0.0);
  w.p_prev.assign(N, 0.0);
  w.eta.assign(eta, eta + N);
  for (int c = 0; c < nc; ++c) {
    Contact k;
    for (int j = 0; j < 4; ++j) {
      k.state[j] = state[4 * c + j];
      k.coef[j] = coef[4 * c + j];
      k.gamma[j] = gamma[4 * c + j];
    }
    for (int d = 0; d < 3; ++d) {
      k.normal[d] = normal[3 * c + d];
      k.affine[d] = affine[3 * c + d];
      k.corr[d] = corr[3 * c + d];
    }
    k.mu = mu[c];
    w.contacts.push_back(k);
  }
  recolor(w);
  std::vector<double> X(x, x + N), dual(N, 0.0), dcur(N, 0.0), dprev(N, 0.0);
  double primal = 0.0;
  contact_update(w, X, dual, dcur, dprev, &primal);
  std::copy(w.p.begin(), w.p.end(), p_out);
  std::copy(w.eta.begin(), w.eta.end(), eta);
  for (int c = 0; c < nc; ++c)
    for (int d = 0; d < 3; ++d) corr[3 * c + d] = w.contacts[c].corr[d];
  for (size_t g = 0; g < w.color_groups.size(); ++g)
    for (int c : w.color_groups[g]) color_out[c] = static_cast<int>(g) + 1;
}

void oracle_set_state(void* wp, int idx, const double* x, const double* v) {
  World& w = *static_cast<World*>(wp);
  auto& c = w.cloths[idx];
  for (int i = 0; i < c.nv; ++i)
    for (int d = 0; d < 3; ++d) {
      c.x[3 * c.perm[i] + d] = x[3 * i + d];
      c.v[3 * c.perm[i] + d] = v[3 * i + d];
    }
}

}  // extern "C"
