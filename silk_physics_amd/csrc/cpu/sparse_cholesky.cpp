#include "cpu/sparse_cholesky.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace silk::amd::cpu {

namespace {

// Nested dissection on positions: split the vertex set at the median of its
// longest extent, take the vertices of the first half that touch the second
// as the separator, order both halves recursively and the separator last.
void dissect(const CsrMatrix& A, const std::vector<float>& pos, std::vector<int> set,
             std::vector<int>& side, int stamp, std::vector<int>& order) {
  constexpr size_t kLeaf = 48;
  if (set.size() <= kLeaf) {
    order.insert(order.end(), set.begin(), set.end());
    return;
  }
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  for (int v : set)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], pos[3 * v + d]);
      hi[d] = std::max(hi[d], pos[3 * v + d]);
    }
  int axis = 0;
  for (int d = 1; d < 3; ++d)
    if (hi[d] - lo[d] > hi[axis] - lo[axis]) axis = d;
  const size_t half = set.size() / 2;
  std::nth_element(set.begin(), set.begin() + static_cast<long>(half), set.end(),
                   [&](int a, int b) {
                     const float pa = pos[3 * a + axis], pb = pos[3 * b + axis];
                     return pa < pb || (pa == pb && a < b);
                   });
  // side[v] == stamp: v is in the second half of this split.
  for (size_t i = half; i < set.size(); ++i) side[set[i]] = stamp;
  std::vector<int> first, sep;
  first.reserve(half);
  for (size_t i = 0; i < half; ++i) {
    const int v = set[i];
    bool touches = false;
    for (int p = A.row_ptr[v]; p < A.row_ptr[v + 1] && !touches; ++p)
      touches = side[A.col[p]] == stamp;
    (touches ? sep : first).push_back(v);
  }
  std::vector<int> second(set.begin() + static_cast<long>(half), set.end());
  set.clear();
  set.shrink_to_fit();
  dissect(A, pos, std::move(first), side, 2 * stamp, order);
  dissect(A, pos, std::move(second), side, 2 * stamp + 1, order);
  order.insert(order.end(), sep.begin(), sep.end());
}

}  // namespace

void SparseCholesky::analyse(const CsrMatrix& A, const std::vector<float>& pos) {
  n_ = A.n;
  std::vector<int> all(n_);
  std::iota(all.begin(), all.end(), 0);
  // Stamps grow by doubling per level; 64-bit would be needed only past
  // 2^30 recursion paths, which a depth of log2(n / 48) never reaches.
  std::vector<int> side(n_, 0);
  perm_.clear();
  perm_.reserve(n_);
  dissect(A, pos, std::move(all), side, 1, perm_);
  iperm_.assign(n_, 0);
  for (int i = 0; i < n_; ++i) iperm_[perm_[i]] = i;

  // Lower part of A by rows in the new numbering.
  ap_.assign(n_ + 1, 0);
  acol_.clear();
  aval_.clear();
  std::vector<std::pair<int, int64_t>> row;
  for (int k = 0; k < n_; ++k) {
    const int o = perm_[k];
    row.clear();
    for (int p = A.row_ptr[o]; p < A.row_ptr[o + 1]; ++p) {
      const int j = iperm_[A.col[p]];
      if (j <= k) row.emplace_back(j, p);
    }
    std::sort(row.begin(), row.end());
    for (const auto& [j, p] : row) {
      acol_.push_back(j);
      aval_.push_back(p);
    }
    ap_[k + 1] = static_cast<int64_t>(acol_.size());
  }

  // Elimination tree (path-compressed ancestors).
  parent_.assign(n_, -1);
  std::vector<int> anc(n_, -1);
  for (int k = 0; k < n_; ++k)
    for (int64_t p = ap_[k]; p < ap_[k + 1]; ++p) {
      int i = acol_[p];
      while (i != -1 && i < k) {
        const int next = anc[i];
        anc[i] = k;
        if (next == -1) parent_[i] = k;
        i = next;
      }
    }

  // Row patterns of L: the etree reach of A's row entries.
  std::vector<int> mark(n_, -1);
  rp_.assign(n_ + 1, 0);
  rcol_.clear();
  std::vector<int> colcount(n_, 1);  // diagonal
  std::vector<int> pat;
  for (int k = 0; k < n_; ++k) {
    pat.clear();
    mark[k] = k;
    for (int64_t p = ap_[k]; p < ap_[k + 1]; ++p)
      for (int i = acol_[p]; i != -1 && mark[i] != k; i = parent_[i]) {
        mark[i] = k;
        pat.push_back(i);
      }
    std::sort(pat.begin(), pat.end());
    for (int j : pat) {
      rcol_.push_back(j);
      ++colcount[j];
    }
    rp_[k + 1] = static_cast<int64_t>(rcol_.size());
  }
  // Columns: diagonal first, then rows ascending (rows are visited in order).
  lp_.assign(n_ + 1, 0);
  for (int j = 0; j < n_; ++j) lp_[j + 1] = lp_[j] + colcount[j];
  li_.assign(static_cast<size_t>(lp_[n_]), 0);
  lx_.assign(static_cast<size_t>(lp_[n_]), 0.0);
  rslot_.assign(rcol_.size(), 0);
  std::vector<int64_t> next(n_);
  for (int j = 0; j < n_; ++j) {
    li_[static_cast<size_t>(lp_[j])] = j;
    next[j] = lp_[j] + 1;
  }
  for (int k = 0; k < n_; ++k)
    for (int64_t p = rp_[k]; p < rp_[k + 1]; ++p) {
      const int j = rcol_[p];
      li_[static_cast<size_t>(next[j])] = k;
      rslot_[p] = next[j]++;
    }
  work_.assign(n_, 0.0);
}

bool SparseCholesky::factor_rows(const CsrMatrix& A, const std::vector<double>& extra,
                                 const std::vector<char>* only) {
  std::vector<double>& x = work_;
  rows_last_ = 0;
  for (int k = 0; k < n_; ++k) {
    if (only != nullptr && !(*only)[k]) continue;
    ++rows_last_;
    for (int64_t p = ap_[k]; p < ap_[k + 1]; ++p)
      x[acol_[p]] += A.val[static_cast<size_t>(aval_[p])];
    double d = x[k] + extra[perm_[k]];
    x[k] = 0.0;
    for (int64_t p = rp_[k]; p < rp_[k + 1]; ++p) {
      const int j = rcol_[p];
      const int64_t c0 = lp_[j];
      const double lkj = x[j] / lx_[static_cast<size_t>(c0)];
      x[j] = 0.0;
      // Entries of column j above row k (rows ascending after the diagonal).
      for (int64_t q = c0 + 1; q < lp_[j + 1]; ++q) {
        const int i = li_[static_cast<size_t>(q)];
        if (i >= k) break;
        x[i] -= lx_[static_cast<size_t>(q)] * lkj;
      }
      d -= lkj * lkj;
      lx_[static_cast<size_t>(rslot_[p])] = lkj;
    }
    if (!(d > 0.0)) return false;
    lx_[static_cast<size_t>(lp_[k])] = std::sqrt(d);
  }
  return true;
}

bool SparseCholesky::factor(const CsrMatrix& A, const std::vector<double>& extra) {
  return factor_rows(A, extra, nullptr);
}

bool SparseCholesky::refactor(const CsrMatrix& A, const std::vector<double>& old_extra,
                              const std::vector<double>& extra) {
  std::vector<char> only(n_, 0);
  bool any = false;
  for (int v = 0; v < n_; ++v) {
    if (old_extra[v] == extra[v]) continue;
    any = true;
    for (int k = iperm_[v]; k != -1 && !only[k]; k = parent_[k]) only[k] = 1;
  }
  if (!any) {
    rows_last_ = 0;
    return true;
  }
  return factor_rows(A, extra, &only);
}

void SparseCholesky::solve3(const double* b, double* x) const {
  std::vector<double> y(3 * static_cast<size_t>(n_));
  for (int k = 0; k < n_; ++k)
    for (int d = 0; d < 3; ++d) y[3 * k + d] = b[3 * perm_[k] + d];
  // L y = b by columns.
  for (int j = 0; j < n_; ++j) {
    const double dj = lx_[static_cast<size_t>(lp_[j])];
    const double y0 = y[3 * j] / dj, y1 = y[3 * j + 1] / dj, y2 = y[3 * j + 2] / dj;
    y[3 * j] = y0;
    y[3 * j + 1] = y1;
    y[3 * j + 2] = y2;
    for (int64_t q = lp_[j] + 1; q < lp_[j + 1]; ++q) {
      const int i = li_[static_cast<size_t>(q)];
      const double l = lx_[static_cast<size_t>(q)];
      y[3 * i] -= l * y0;
      y[3 * i + 1] -= l * y1;
      y[3 * i + 2] -= l * y2;
    }
  }
  // L^T x = y by columns (dot products).
  for (int j = n_ - 1; j >= 0; --j) {
    double s0 = y[3 * j], s1 = y[3 * j + 1], s2 = y[3 * j + 2];
    for (int64_t q = lp_[j] + 1; q < lp_[j + 1]; ++q) {
      const int i = li_[static_cast<size_t>(q)];
      const double l = lx_[static_cast<size_t>(q)];
      s0 -= l * y[3 * i];
      s1 -= l * y[3 * i + 1];
      s2 -= l * y[3 * i + 2];
    }
    const double dj = lx_[static_cast<size_t>(lp_[j])];
    y[3 * j] = s0 / dj;
    y[3 * j + 1] = s1 / dj;
    y[3 * j + 2] = s2 / dj;
  }
  for (int k = 0; k < n_; ++k)
    for (int d = 0; d < 3; ++d) x[3 * perm_[k] + d] = y[3 * k + d];
}

}  // namespace silk::amd::cpu
