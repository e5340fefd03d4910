// Sparse Cholesky factor of the vertex-scalar PD system of the CPU backend.
//
// The reference factors the 3N system H = M/h^2 + k JWJ + kb CWC + pins with
// CHOLMOD and applies barrier terms as a rank update of that factor
// (cloth_solver_context.cpp:21-73, cloth_solver_utils.cpp:144-198).  Every
// term of H is a vertex scalar times I3, so one N x N factor serves the three
// coordinates.  Order: geometric nested dissection of the rest positions
// (median planes, the part-side vertices adjacent to the other part as the
// separator, ordered last).  Factor: up-looking, row by row, L stored by
// columns with the positions of every row's entries fixed by the symbolic
// analysis.  Barrier terms add to the diagonal; a change of the diagonal at
// vertex i changes only the rows of i and its elimination-tree ancestors, so
// `refactor` recomputes those rows alone.
#pragma once

#include <cstdint>
#include <vector>

#include "host/geometry.hpp"

namespace silk::amd::cpu {

class SparseCholesky {
 public:
  /// Symbolic analysis of A's pattern (symmetric CSR, diagonal included) in a
  /// nested-dissection order of the vertices at `pos` (3 floats each).
  void analyse(const CsrMatrix& A, const std::vector<float>& pos);
  /// Numeric factor of A + diag(extra) (extra: one value per vertex).  False
  /// when a pivot is not positive.
  bool factor(const CsrMatrix& A, const std::vector<double>& extra);
  /// Re-factors after the diagonal changed from `old_extra` to `extra`: only
  /// rows whose values depend on a changed entry are recomputed.
  bool refactor(const CsrMatrix& A, const std::vector<double>& old_extra,
                const std::vector<double>& extra);
  /// x = (A + diag(extra))^-1 b for three interleaved right-hand sides
  /// (x, y, z of every vertex, original numbering).
  void solve3(const double* b, double* x) const;

  int n() const { return n_; }
  int64_t nnz() const { return static_cast<int64_t>(li_.size()); }
  int rows_recomputed() const { return rows_last_; }

 private:
  bool factor_rows(const CsrMatrix& A, const std::vector<double>& extra,
                   const std::vector<char>* only);

  int n_ = 0;
  std::vector<int> perm_, iperm_;  // perm_[new] = old
  std::vector<int> parent_;        // elimination tree (new numbering)
  // L by columns: lp_[j] .. lp_[j+1], row indices ascending, diagonal first.
  std::vector<int64_t> lp_;
  std::vector<int> li_;
  std::vector<double> lx_;
  // Row k's off-diagonal entries: columns ascending and their slots in lx_.
  std::vector<int64_t> rp_;
  std::vector<int> rcol_;
  std::vector<int64_t> rslot_;
  // A's entries of row k in the new numbering with column <= k, as
  // (column, index into A.val).
  std::vector<int64_t> ap_;
  std::vector<int> acol_;
  std::vector<int64_t> aval_;
  mutable std::vector<double> work_;
  int rows_last_ = 0;
};

}  // namespace silk::amd::cpu
