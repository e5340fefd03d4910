// Oracle: static cloth operators in the libigl / reference formulation.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <algorithm>
#include <cmath>

#include "oracle.hpp"

namespace oracle {

namespace {
V3 vrow(const std::vector<double>& V, int i) {
  return {V[3 * i], V[3 * i + 1], V[3 * i + 2]};
}
double norm(const V3& a) { return std::sqrt(dot(a, a)); }
}  // namespace

// cloth_assembly_l2_cache.cpp:27-69: vec(F) = J x, J = kron(B, I3),
// B = (D dX^-1)^T with the 2-D chart x axis along e0.
static std::array<double, 54> jacobian(const V3& v0, const V3& v1, const V3& v2) {
  V3 e0 = v1 - v0, e1 = v2 - v0;
  V3 n = cross(e0, e1);
  V3 bx = (1.0 / norm(e0)) * e0;
  V3 by0 = cross(n, e0);
  V3 by = (1.0 / norm(by0)) * by0;
  double dX[2][2] = {{dot(bx, e0), dot(bx, e1)}, {0.0, dot(by, e1)}};
  double det = dX[0][0] * dX[1][1] - dX[0][1] * dX[1][0];
  double inv[2][2] = {{dX[1][1] / det, -dX[0][1] / det},
                      {-dX[1][0] / det, dX[0][0] / det}};
  const double D[3][2] = {{-1, -1}, {1, 0}, {0, 1}};
  double DI[3][2];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 2; ++c)
      DI[r][c] = D[r][0] * inv[0][c] + D[r][1] * inv[1][c];
  // B(a, b) = DI(b, a); J(3a + r, 3b + c) = B(a, b) delta(r, c).
  std::array<double, 54> J{};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 3; ++b)
      for (int r = 0; r < 3; ++r) J[9 * (3 * a + r) + (3 * b + r)] = DI[b][a];
  return J;
}

L2Cache build_l2(const std::vector<double>& V, const std::vector<int>& F) {
  const int nv = static_cast<int>(V.size() / 3);
  const int nf = static_cast<int>(F.size() / 3);
  L2Cache l2;
  l2.mass.assign(nv, 0.0);
  l2.area.assign(nf, 0.0);
  l2.cot.assign(nv, {});
  l2.jop.resize(nf);
  for (int f = 0; f < nf; ++f) {
    const int id[3] = {F[3 * f], F[3 * f + 1], F[3 * f + 2]};
    const V3 p[3] = {vrow(V, id[0]), vrow(V, id[1]), vrow(V, id[2])};
    const double dblA = norm(cross(p[1] - p[0], p[2] - p[0]));
    l2.area[f] = 0.5 * dblA;
    // l[i]: length of the edge opposite corner i (libigl edge_lengths).
    double l[3], sq[3];
    for (int i = 0; i < 3; ++i) {
      V3 e = p[(i + 1) % 3] - p[(i + 2) % 3];
      sq[i] = dot(e, e);
      l[i] = std::sqrt(sq[i]);
    }
    // Cotangent entries: half cot of the angle at corner i from squared
    // lengths (libigl cotmatrix_entries).
    for (int i = 0; i < 3; ++i) {
      const double c = (sq[(i + 1) % 3] + sq[(i + 2) % 3] - sq[i]) / dblA / 4.0;
      const int a = id[(i + 1) % 3], b = id[(i + 2) % 3];
      l2.cot[a][b] += c;
      l2.cot[b][a] += c;
      l2.cot[a][a] -= c;
      l2.cot[b][b] -= c;
    }
    // Voronoi mass via circumcentre barycentrics (libigl massmatrix).
    double cosv[3], bary[3], bsum = 0.0;
    for (int i = 0; i < 3; ++i) {
      const int j = (i + 1) % 3, k = (i + 2) % 3;
      cosv[i] = (sq[j] + sq[k] - sq[i]) / (2.0 * l[j] * l[k]);
      bary[i] = cosv[i] * l[i];
      bsum += bary[i];
    }
    double part[3], quad[3];
    for (int i = 0; i < 3; ++i) part[i] = bary[i] / bsum * dblA * 0.5;
    for (int i = 0; i < 3; ++i)
      quad[i] = 0.5 * (part[(i + 1) % 3] + part[(i + 2) % 3]);
    for (int o = 0; o < 3; ++o) {
      if (cosv[o] < 0.0) {
        for (int i = 0; i < 3; ++i) quad[i] = (i == o) ? 0.25 * dblA : 0.125 * dblA;
      }
    }
    for (int i = 0; i < 3; ++i) l2.mass[id[i]] += quad[i];
    l2.jop[f] = jacobian(p[0], p[1], p[2]);
  }
  // CWC = C^T W C.
  l2.cwc.assign(nv, {});
  for (int k = 0; k < nv; ++k) {
    const double w = 1.0 / l2.mass[k];
    for (const auto& [i, ci] : l2.cot[k])
      for (const auto& [j, cj] : l2.cot[k]) l2.cwc[i][j] += ci * w * cj;
  }
  l2.c0.assign(3 * nv, 0.0);
  for (int i = 0; i < nv; ++i)
    for (const auto& [j, v] : l2.cwc[i])
      for (int d = 0; d < 3; ++d) l2.c0[3 * i + d] += v * V[3 * j + d];
  return l2;
}

void Bsr3::add(int i, int j, const double* blk9, double s) {
  auto& b = rows[i][j];
  for (int e = 0; e < 9; ++e) b[e] += s * blk9[e];
}

void Bsr3::mul(const double* x, double* y) const {
  for (int i = 0; i < n; ++i) {
    double acc[3] = {0, 0, 0};
    for (const auto& [j, b] : rows[i])
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) acc[r] += b[3 * r + c] * x[3 * j + c];
    for (int r = 0; r < 3; ++r) y[3 * i + r] = acc[r];
  }
}

Bsr3 build_operator(const L2Cache& l2, const std::vector<int>& F, double k,
                    double kb) {
  Bsr3 A;
  const int nv = static_cast<int>(l2.mass.size());
  A.n = nv;
  A.rows.assign(nv, {});
  const int nf = static_cast<int>(F.size() / 3);
  for (int f = 0; f < nf; ++f) {
    const auto& J = l2.jop[f];
    double AA[9][9] = {};
    for (int r = 0; r < 9; ++r)
      for (int c = 0; c < 9; ++c) {
        double s = 0.0;
        for (int m = 0; m < 6; ++m) s += J[9 * m + r] * J[9 * m + c];
        AA[r][c] = l2.area[f] * s;
      }
    for (int vi = 0; vi < 3; ++vi)
      for (int vj = 0; vj < 3; ++vj) {
        double blk[9];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) blk[3 * r + c] = AA[3 * vi + r][3 * vj + c];
        A.add(F[3 * f + vi], F[3 * f + vj], blk, k);
      }
  }
  const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int i = 0; i < nv; ++i)
    for (const auto& [j, v] : l2.cwc[i]) A.add(i, j, I3, kb * v);
  return A;
}

double sym3_min_eig(const double a_in[9]) {
  double a[3][3];
  for (int i = 0; i < 9; ++i) a[i / 3][i % 3] = a_in[i];
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    if (off < 1e-30) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (std::fabs(a[p][q]) < 1e-300) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) /
                         (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
      }
  }
  return std::min({a[0][0], a[1][1], a[2][2]});
}

}  // namespace oracle
