// Oracle: unit-level C entry points for the parity tests (ctypes).
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "oracle.hpp"

using namespace oracle;

extern "C" {

/// Voronoi mass (nv) and face area (nf) of a mesh.
void oracle_mass_area(const double* V, int nv, const int* F, int nf,
                      double* mass, double* area) {
  std::vector<double> Vv(V, V + 3 * nv);
  std::vector<int> Fv(F, F + 3 * nf);
  L2Cache l2 = build_l2(Vv, Fv);
  std::copy(l2.mass.begin(), l2.mass.end(), mass);
  std::copy(l2.area.begin(), l2.area.end(), area);
}

/// 6x9 Jacobians (row-major, 54 per face).
void oracle_jacobians(const double* V, int nv, const int* F, int nf,
                      double* jop) {
  std::vector<double> Vv(V, V + 3 * nv);
  std::vector<int> Fv(F, F + 3 * nf);
  L2Cache l2 = build_l2(Vv, Fv);
  for (int f = 0; f < nf; ++f) std::copy(l2.jop[f].begin(), l2.jop[f].end(), jop + 54 * f);
}

/// Dense 3nv x 3nv operator k JWJ + kb CWC (x) I3 and the bending rhs
/// kb * CWC * V (3nv).  For small meshes only.
void oracle_operator_dense(const double* V, int nv, const int* F, int nf,
                           double k, double kb, double* A, double* bend_rhs) {
  std::vector<double> Vv(V, V + 3 * nv);
  std::vector<int> Fv(F, F + 3 * nf);
  L2Cache l2 = build_l2(Vv, Fv);
  Bsr3 M = build_operator(l2, Fv, k, kb);
  const size_t n3 = 3 * static_cast<size_t>(nv);
  std::fill(A, A + n3 * n3, 0.0);
  for (int i = 0; i < nv; ++i)
    for (const auto& [j, b] : M.rows[i])
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) A[(3 * i + r) * n3 + 3 * j + c] = b[3 * r + c];
  if (bend_rhs)
    for (size_t i = 0; i < n3; ++i) bend_rhs[i] = kb * l2.c0[i];
}

void oracle_project_stretch(const double* y, double lo, double hi, double* out) {
  project_stretch(y, lo, hi, out);
}

/// One ADMM aux update for every face (cloth_admm_helper.cu:142-243) given the
/// rest mesh (Jacobians), penalty rho = k, current x (3nv), u and z (6nf, in
/// and out).  sums[3] = primal, ||W Sx||^2, ||W z||^2; dual/dcur/dprev (3nv).
void oracle_local_step(const double* V, int nv, const int* F, int nf, double k,
                       const double* x, double* u, double* z, double* sums,
                       double* dual, double* dcur, double* dprev) {
  std::vector<double> Vv(V, V + 3 * nv);
  std::vector<int> Fv(F, F + 3 * nf);
  L2Cache l2 = build_l2(Vv, Fv);
  std::fill(sums, sums + 3, 0.0);
  std::fill(dual, dual + 3 * nv, 0.0);
  std::fill(dcur, dcur + 3 * nv, 0.0);
  std::fill(dprev, dprev + 3 * nv, 0.0);
  const double rho = k;
  for (int f = 0; f < nf; ++f) {
    const auto& J = l2.jop[f];
    const double w = std::sqrt(l2.area[f]);
    double xl[9], Sx[6], y[6], pr[6], zn[6];
    for (int c = 0; c < 9; ++c) xl[c] = x[3 * F[3 * f + c / 3] + c % 3];
    for (int r = 0; r < 6; ++r) {
      Sx[r] = 0.0;
      for (int c = 0; c < 9; ++c) Sx[r] += J[9 * r + c] * xl[c];
      y[r] = Sx[r] + u[6 * f + r] / (rho * w);
    }
    project_stretch(y, 0.9, 1.1, pr);
    const double rr = k / (k + rho);
    for (int r = 0; r < 6; ++r) zn[r] = (1 - rr) * y[r] + rr * pr[r];
    for (int r = 0; r < 6; ++r) {
      sums[0] += (w * (Sx[r] - zn[r])) * (w * (Sx[r] - zn[r]));
      sums[1] += (w * Sx[r]) * (w * Sx[r]);
      sums[2] += (w * zn[r]) * (w * zn[r]);
    }
    for (int c = 0; c < 9; ++c) {
      double a = 0, b = 0;
      for (int r = 0; r < 6; ++r) {
        a += J[9 * r + c] * zn[r];
        b += J[9 * r + c] * z[6 * f + r];
      }
      const int gi = 3 * F[3 * f + c / 3] + c % 3;
      dual[gi] += rho * w * w * (a - b);
      dcur[gi] += rho * w * w * a;
      dprev[gi] += rho * w * w * b;
    }
    for (int r = 0; r < 6; ++r) {
      z[6 * f + r] = zn[r];
      u[6 * f + r] = std::clamp(u[6 * f + r] + w * rho * (Sx[r] - zn[r]), -1e8, 1e8);
    }
  }
}

/// MAS preconditioner (reference 96x96 formulation) applied once:
/// z = M^-1 r for the operator of the mesh with the given diagonal (3nv).
/// The mesh must already be in partition order; offsets has nparts+1 ints.
void oracle_mas_apply(const double* V, int nv, const int* F, int nf, double k,
                      double kb, const int* offsets, int nparts,
                      const double* diag, const double* r, double* z,
                      int* level_num) {
  std::vector<double> Vv(V, V + 3 * nv);
  std::vector<int> Fv(F, F + 3 * nf);
  L2Cache l2 = build_l2(Vv, Fv);
  Bsr3 A = build_operator(l2, Fv, k, kb);
  MasOracle M;
  M.build_topology(A, std::vector<int>(offsets, offsets + nparts + 1));
  M.factorize(A, std::vector<double>(diag, diag + 3 * nv));
  M.apply(r, z);
  if (level_num) *level_num = M.level_num;
}

/// Reference PCG on the same system: x (in: initial guess, out: solution).
/// out[0] iterations, out[1] status, out[2] rr0, out[3] rr.
void oracle_pcg(const double* V, int nv, const int* F, int nf, double k,
                double kb, const int* offsets, int nparts, const double* diag,
                const double* b, double* x, double rel_tol, double abs_tol,
                int max_iter, double* out, int criterion) {
  std::vector<double> Vv(V, V + 3 * nv);
  std::vector<int> Fv(F, F + 3 * nf);
  L2Cache l2 = build_l2(Vv, Fv);
  Bsr3 A = build_operator(l2, Fv, k, kb);
  MasOracle M;
  M.build_topology(A, std::vector<int>(offsets, offsets + nparts + 1));
  std::vector<double> dg(diag, diag + 3 * nv);
  M.factorize(A, dg);
  PcgResult r = pcg_solve(A, dg, M, b, x, rel_tol, abs_tol, max_iter, criterion);
  out[0] = r.iters;
  out[1] = r.status;
  out[2] = r.rr0;
  out[3] = r.rr;
}

/// Boxes: 6 floats each (min xyz, max xyz).  Writes up to cap pairs (2 ints
/// each), returns the total count.
int oracle_overlap_pairs(const float* a, int na, const float* b, int nb,
                         int self, int* pairs, int cap) {
  auto load = [](const float* p, int n) {
    std::vector<Box> v(n);
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) {
        v[i].mn[d] = p[6 * i + d];
        v[i].mx[d] = p[6 * i + 3 + d];
      }
    return v;
  };
  auto A = load(a, na);
  auto B = self ? A : load(b, nb);
  auto out = overlap_pairs(A, B, self != 0);
  const int n = static_cast<int>(out.size());
  for (int i = 0; i < std::min(n, cap); ++i) {
    pairs[2 * i] = out[i].first;
    pairs[2 * i + 1] = out[i].second;
  }
  return n;
}

/// FP64 CCD of one query.  p0, p1: 12 doubles (4 points).  Returns 1 on hit,
/// 2 for no hit with a root at t = 0 at the last separation searched, else 0.
int oracle_ccd(int vf, const double* p0, const double* p1, double ms,
               double tolerance, int max_iter, const double* err,
               double max_time, int stack_cap, int zero_mode, double* toi) {
  double a[4][3], b[4][3];
  for (int i = 0; i < 12; ++i) {
    a[i / 3][i % 3] = p0[i];
    b[i / 3][i % 3] = p1[i];
  }
  CcdConfig c;
  c.ms = ms;
  c.tolerance = tolerance;
  c.max_iter = max_iter;
  for (int d = 0; d < 3; ++d) c.err[d] = err[d];
  c.max_time = max_time;
  c.stack_cap = stack_cap;
  c.zero_mode = zero_mode;
  bool zero = false;
  if (ccd_query(vf != 0, a, b, c, toi, &zero)) return 1;
  return zero ? 2 : 0;
}

/// The same query in single precision (the reference CPU CCD as it runs).
int oracle_ccd_f32(int vf, const float* p0, const float* p1, float ms,
                   float tolerance, int max_iter, const float* err,
                   float max_time, float* toi) {
  float a[4][3], b[4][3];
  for (int i = 0; i < 12; ++i) {
    a[i / 3][i % 3] = p0[i];
    b[i / 3][i % 3] = p1[i];
  }
  return ccd_query_f32(vf != 0, a, b, ms, tolerance, max_iter, err, max_time, toi) ? 1 : 0;
}

void oracle_coulomb(const double* us, const double* n, double mu, double* out) {
  coulomb_solve(us, n, mu, out);
}

/// x: 12 doubles (two edges).  out: (u, v).  Returns 1 when not degenerate.
int oracle_ee_uv(const double* x, double eps, double* out) {
  V3 p[4];
  for (int i = 0; i < 4; ++i) p[i] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
  return ee_uv(p[0], p[1], p[2], p[3], eps, &out[0], &out[1]) ? 1 : 0;
}

void oracle_ccd_filter(const double* abs_max, int vf, double* err) {
  ccd_filter(abs_max, vf != 0, err);
}

}  // extern "C"
