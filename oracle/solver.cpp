// Oracle: MAS preconditioner, PCG and the stretch projection in the
// reference formulation (96x96 banks of 3x3 blocks), in double.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "oracle.hpp"

namespace oracle {

namespace {

constexpr int kBank = 32;
constexpr int kDim = 96;

int ctz(uint32_t x) { return __builtin_ctz(x); }
int popc(uint32_t x) { return __builtin_popcount(x); }

// Bank-local connected components by neighbour-mask closure
// (mas_preconditioner.cu:214-236).  masks[s] holds the in-bank neighbours of
// slot s (self included); valid marks real slots.  Returns local ids.
std::vector<int> bank_components(const std::array<uint32_t, kBank>& masks,
                                 const std::array<bool, kBank>& valid,
                                 int* count) {
  uint32_t leads = 0;
  std::array<int, kBank> lead_of{};
  for (int s = 0; s < kBank; ++s) {
    if (!valid[s]) continue;
    uint32_t conn = masks[s];
    uint32_t visited = 1u << s;
    uint32_t todo = conn ^ visited;
    while (todo) {
      int v = ctz(todo);
      conn |= masks[v];
      visited |= 1u << v;
      todo = conn ^ visited;
    }
    lead_of[s] = ctz(conn);
    if (lead_of[s] == s) leads |= 1u << s;
  }
  std::vector<int> id(kBank, -1);
  for (int s = 0; s < kBank; ++s)
    if (valid[s])
      id[s] = popc(leads & static_cast<uint32_t>((1ull << lead_of[s]) - 1ull));
  *count = popc(leads);
  return id;
}

// Dense symmetric inverse by Gauss-Jordan with partial pivoting.
std::vector<double> dense_inverse(std::vector<double> a, int n) {
  std::vector<double> inv(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r * n + c]) > std::fabs(a[piv * n + c])) piv = r;
    if (a[piv * n + c] == 0.0) throw std::runtime_error("singular MAS block");
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[c * n + k], a[piv * n + k]);
        std::swap(inv[c * n + k], inv[piv * n + k]);
      }
    const double ip = 1.0 / a[c * n + c];
    for (int k = 0; k < n; ++k) {
      a[c * n + k] *= ip;
      inv[c * n + k] *= ip;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[r * n + k] -= f * a[c * n + k];
        inv[r * n + k] -= f * inv[c * n + k];
      }
    }
  }
  return inv;
}

}  // namespace

void MasOracle::build_topology(const Bsr3& A,
                               const std::vector<int>& part_offsets) {
  n = A.n;
  const int parts = static_cast<int>(part_offsets.size()) - 1;
  padded = parts * kBank;
  real_to_padded.assign(n, -1);
  padded_to_real.assign(padded, -1);
  for (int p = 0; p < parts; ++p)
    for (int i = part_offsets[p]; i < part_offsets[p + 1]; ++i) {
      real_to_padded[i] = p * kBank + (i - part_offsets[p]);
      padded_to_real[p * kBank + (i - part_offsets[p])] = i;
    }
  map.assign(static_cast<size_t>(n) * 4, -1);
  // Level 0 CCOs (mas_preconditioner.cu:162-237,338-352).
  int total = 0;
  for (int b = 0; b < parts; ++b) {
    std::array<uint32_t, kBank> masks{};
    std::array<bool, kBank> valid{};
    for (int s = 0; s < kBank; ++s) {
      const int i = padded_to_real[b * kBank + s];
      if (i < 0) continue;
      valid[s] = true;
      masks[s] = 1u << s;
      for (const auto& [j, blk] : A.rows[i]) {
        (void)blk;
        const int pj = real_to_padded[j];
        if (pj / kBank == b) masks[s] |= 1u << (pj % kBank);
      }
    }
    int cnt = 0;
    auto ids = bank_components(masks, valid, &cnt);
    for (int s = 0; s < kBank; ++s) {
      const int i = padded_to_real[b * kBank + s];
      if (i >= 0) map[4 * i] = total + ids[s];
    }
    total += cnt;
  }
  cco_nums = {};
  cco_nums[0] = total;
  level_num = 1;
  // Coarser levels (mas_preconditioner.cu:246-332,358-374,434-471).
  for (int lv = 1; lv < 4; ++lv) {
    const int prev = cco_nums[lv - 1];
    const int banks = (prev + kBank - 1) / kBank;
    std::vector<uint32_t> masks(prev, 0);
    for (int c = 0; c < prev; ++c) masks[c] = 1u << (c % kBank);
    for (int i = 0; i < n; ++i)
      for (const auto& [j, blk] : A.rows[i]) {
        (void)blk;
        const int I = map[4 * i + lv - 1], J = map[4 * j + lv - 1];
        if (I / kBank == J / kBank) masks[I] |= 1u << (J % kBank);
      }
    int next = 0;
    std::vector<int> new_id(prev, -1);
    for (int b = 0; b < banks; ++b) {
      std::array<uint32_t, kBank> bm{};
      std::array<bool, kBank> valid{};
      for (int s = 0; s < kBank; ++s) {
        const int c = b * kBank + s;
        if (c >= prev) continue;
        valid[s] = true;
        bm[s] = masks[c];
      }
      int cnt = 0;
      auto ids = bank_components(bm, valid, &cnt);
      for (int s = 0; s < kBank; ++s)
        if (valid[s]) new_id[b * kBank + s] = next + ids[s];
      next += cnt;
    }
    for (int i = 0; i < n; ++i) map[4 * i + lv] = new_id[map[4 * i + lv - 1]];
    if (next == prev) break;
    cco_nums[lv] = next;
    level_num = lv + 1;
  }
  mat_nums.assign(level_num + 1, 0);
  mat_offsets.assign(level_num + 1, 0);
  mat_nums[0] = padded / kBank;
  for (int l = 0; l < level_num; ++l)
    mat_nums[l + 1] = (cco_nums[l] + kBank - 1) / kBank;
  for (int l = 1; l <= level_num; ++l)
    mat_offsets[l] = mat_offsets[l - 1] + mat_nums[l - 1];
}

void MasOracle::factorize(const Bsr3& A, const std::vector<double>& diag) {
  const int total = mat_offsets[level_num] + mat_nums[level_num];
  std::vector<std::vector<double>> dense(
      total, std::vector<double>(static_cast<size_t>(kDim) * kDim, 0.0));
  auto put = [&](int mat, int row, int col, double v) {
    dense[mat][static_cast<size_t>(row) * kDim + col] += v;
  };
  for (int i = 0; i < n; ++i) {
    const int pi = real_to_padded[i];
    for (const auto& [j, blk] : A.rows[i]) {
      const int pj = real_to_padded[j];
      if (pi / kBank == pj / kBank)
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c)
            put(pi / kBank, 3 * (pi % kBank) + r, 3 * (pj % kBank) + c,
                blk[3 * r + c]);
      for (int l = 0; l < level_num; ++l) {
        const int I = map[4 * i + l], J = map[4 * j + l];
        if (I / kBank != J / kBank) continue;
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c)
            put(mat_offsets[l + 1] + I / kBank, 3 * (I % kBank) + r,
                3 * (J % kBank) + c, blk[3 * r + c]);
      }
    }
    for (int d = 0; d < 3; ++d) {
      put(pi / kBank, 3 * (pi % kBank) + d, 3 * (pi % kBank) + d, diag[3 * i + d]);
      for (int l = 0; l < level_num; ++l) {
        const int I = map[4 * i + l];
        put(mat_offsets[l + 1] + I / kBank, 3 * (I % kBank) + d,
            3 * (I % kBank) + d, diag[3 * i + d]);
      }
    }
  }
  inv.assign(total, {});
  for (int m = 0; m < total; ++m) {
    for (int d = 0; d < kDim; ++d)
      if (dense[m][static_cast<size_t>(d) * kDim + d] == 0.0)
        dense[m][static_cast<size_t>(d) * kDim + d] = 1.0;  // padding slot
    inv[m] = dense_inverse(dense[m], kDim);
  }
}

void MasOracle::apply(const double* r, double* z) const {
  std::fill(z, z + 3 * static_cast<size_t>(n), 0.0);
  // level 0
  for (int m = 0; m < mat_nums[0]; ++m) {
    double rb[kDim] = {};
    for (int s = 0; s < kBank; ++s) {
      const int i = padded_to_real[m * kBank + s];
      if (i >= 0)
        for (int d = 0; d < 3; ++d) rb[3 * s + d] = r[3 * i + d];
    }
    for (int s = 0; s < kBank; ++s) {
      const int i = padded_to_real[m * kBank + s];
      if (i < 0) continue;
      for (int d = 0; d < 3; ++d) {
        double acc = 0.0;
        for (int k = 0; k < kDim; ++k)
          acc += inv[m][static_cast<size_t>(3 * s + d) * kDim + k] * rb[k];
        z[3 * i + d] += acc;
      }
    }
  }
  for (int l = 0; l < level_num; ++l) {
    const int nodes = cco_nums[l];
    std::vector<double> rl(3 * static_cast<size_t>(mat_nums[l + 1]) * kBank, 0.0);
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) rl[3 * map[4 * i + l] + d] += r[3 * i + d];
    std::vector<double> zl(rl.size(), 0.0);
    for (int m = 0; m < mat_nums[l + 1]; ++m) {
      const auto& M = inv[mat_offsets[l + 1] + m];
      for (int row = 0; row < kDim; ++row) {
        double acc = 0.0;
        for (int k = 0; k < kDim; ++k)
          acc += M[static_cast<size_t>(row) * kDim + k] * rl[m * kDim + k];
        zl[m * kDim + row] = acc;
      }
    }
    (void)nodes;
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) z[3 * i + d] += zl[3 * map[4 * i + l] + d];
  }
}

PcgResult pcg_solve(const Bsr3& A, const std::vector<double>& diag,
                    const MasOracle& M, const double* b, double* x,
                    double rel_tol, double abs_tol, int max_iter,
                    int criterion) {
  const size_t N = 3 * static_cast<size_t>(A.n);
  std::vector<double> r(N), z(N), p(N), Ap(N);
  auto spmv = [&](const double* src, double* dst) {
    A.mul(src, dst);
    for (size_t i = 0; i < N; ++i) dst[i] += diag[i] * src[i];
  };
  auto dotv = [&](const std::vector<double>& u, const std::vector<double>& v) {
    double s = 0.0;
    for (size_t i = 0; i < N; ++i) s += u[i] * v[i];
    return s;
  };
  PcgResult res;
  spmv(x, r.data());
  for (size_t i = 0; i < N; ++i) r[i] = b[i] - r[i];
  M.apply(r.data(), z.data());
  p = z;
  double rz = dotv(r, z);
  double rr = dotv(r, r);
  if (!std::isfinite(rz)) {
    res.status = 3;
    return res;
  }
  res.rr0 = rr;
  // Reference of the relative test: rr0 (mas_cg_solver.cu:87-90), or |b|^2
  // for the right-hand-side criterion, under which a warm start that already
  // meets the tolerance takes no iteration.
  double ref2 = rr;
  if (criterion >= 1) {  // 2 (true residual) is 1 in FP64: no drift
    ref2 = 0.0;
    for (size_t i = 0; i < N; ++i) ref2 += b[i] * b[i];
    if (rr <= abs_tol * abs_tol || rr <= rel_tol * rel_tol * ref2) {
      res.status = rr <= abs_tol * abs_tol ? 0 : 1;
      res.rr = rr;
      return res;
    }
  }
  int iter = 0;
  for (;;) {
    spmv(p.data(), Ap.data());
    const double pAp = dotv(p, Ap);
    const double alpha = std::fabs(pAp) < 1e-20 ? 0.0 : rz / pAp;
    for (size_t i = 0; i < N; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
    }
    M.apply(r.data(), z.data());
    const double rz_old = rz;
    rz = dotv(r, z);
    rr = dotv(r, r);
    ++iter;
    int status = -1;
    if (iter >= max_iter) status = 2;
    else if (rr <= abs_tol * abs_tol) status = 0;
    else if (rr <= rel_tol * rel_tol * ref2) status = 1;
    if (status >= 0) {
      res.status = status;
      break;
    }
    const double beta = std::fabs(rz_old) < 1e-20 ? 0.0 : rz / rz_old;
    for (size_t i = 0; i < N; ++i) p[i] = z[i] + beta * p[i];
  }
  res.iters = iter;
  res.rr = rr;
  return res;
}

void project_stretch(const double y[6], double lo, double hi, double out[6]) {
  V3 c0 = {y[0], y[1], y[2]}, c1 = {y[3], y[4], y[5]};
  // One-sided Jacobi: rotate the columns until orthogonal.
  double v[2][2] = {{1.0, 0.0}, {0.0, 1.0}};  // columns v1, v2
  const double a = dot(c0, c0), b = dot(c1, c1), g = dot(c0, c1);
  if (g != 0.0) {
    const double zeta = (b - a) / (2.0 * g);
    const double t = (zeta >= 0 ? 1.0 : -1.0) /
                     (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
    const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
    V3 n0 = c * c0 - s * c1, n1 = s * c0 + c * c1;
    c0 = n0;
    c1 = n1;
    v[0][0] = c;  v[1][0] = -s;  // v1 = (c, -s)
    v[0][1] = s;  v[1][1] = c;   // v2 = (s, c)
  }
  double s0 = std::sqrt(dot(c0, c0)), s1 = std::sqrt(dot(c1, c1));
  // Order by singular value (largest first).
  int L = 0, S = 1;
  if (s1 > s0) std::swap(L, S);
  V3 col[2] = {c0, c1};
  double sig[2] = {s0, s1};
  V3 uL, uS;
  const double tiny = 1e-12;
  if (sig[L] > tiny) {
    uL = (1.0 / sig[L]) * col[L];
  } else {
    uL = {1.0, 0.0, 0.0};
    L = 0;
    S = 1;
    v[0][0] = 1.0; v[1][0] = 0.0; v[0][1] = 0.0; v[1][1] = 1.0;
    sig[0] = sig[1] = 0.0;
  }
  if (sig[S] > 1e-6 * sig[L] && sig[S] > tiny) {
    uS = (1.0 / sig[S]) * col[S];
  } else {
    const double ax = std::fabs(uL[0]), ay = std::fabs(uL[1]), az = std::fabs(uL[2]);
    int axis = 0;
    double amin = ax;
    if (ay < amin) { axis = 1; amin = ay; }
    if (az < amin) axis = 2;
    V3 e = {axis == 0 ? 1.0 : 0.0, axis == 1 ? 1.0 : 0.0, axis == 2 ? 1.0 : 0.0};
    V3 t = e - dot(e, uL) * uL;
    uS = (1.0 / std::sqrt(dot(t, t))) * t;
  }
  const double cL = std::clamp(sig[L], lo, hi), cS = std::clamp(sig[S], lo, hi);
  // proj = cL uL vL^T + cS uS vS^T ; column k of proj = cL vL[k] uL + cS vS[k] uS
  for (int k = 0; k < 2; ++k) {
    V3 colk = (cL * v[k][L]) * uL + (cS * v[k][S]) * uS;
    for (int d = 0; d < 3; ++d) out[3 * k + d] = colk[d];
  }
}

}  // namespace oracle
