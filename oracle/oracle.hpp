// FP64 CPU oracle of the silk GPU projective-dynamics step.
//
// TEST INFRASTRUCTURE ONLY.  This code restates the reference algorithm of
// silk/src/backend/cuda (main_loop.cu, solver/*.cu, collision/*.cu) and the
// reference CPU CCD (backend/cpu/collision/interval_root_finder.cpp) in double
// precision, in the reference's own formulation (3x3 BSR operator, 96x96 MAS
// banks, SVD-based stretch projection).  It shares no source with the
// product library; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as a checker / reported baseline.
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace oracle {

using V3 = std::array<double, 3>;

inline V3 operator+(const V3& a, const V3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 operator-(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 operator*(double s, const V3& a) { return {s * a[0], s * a[1], s * a[2]}; }
inline double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 cross(const V3& a, const V3& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}

// ------------------------------------------------------------------ geometry

/// libigl-style static operators for one mesh (cloth_assembly_l2_cache.cpp).
struct L2Cache {
  std::vector<double> mass;  // Voronoi, per vertex
  std::vector<double> area;  // per face
  std::vector<std::array<double, 54>> jop;  // 6x9 row-major per face
  // vertex-level sparse maps (row -> col -> value)
  std::vector<std::map<int, double>> cot;   // cotangent Laplacian C
  std::vector<std::map<int, double>> cwc;   // C^T W C
  std::vector<double> c0;                   // CWC * V (3*nv)
};
L2Cache build_l2(const std::vector<double>& V, const std::vector<int>& F);

/// 3x3-block sparse matrix (BSR) with a std::map per block row.
struct Bsr3 {
  int n = 0;
  std::vector<std::map<int, std::array<double, 9>>> rows;
  void add(int i, int j, const double* blk9, double s);
  void mul(const double* x, double* y) const;  // y = A x (3n)
};

/// A = k JWJ + kb (CWC (x) I3) (cloth_assembly_l1_cache.cu:264-278).
Bsr3 build_operator(const L2Cache& l2, const std::vector<int>& F, double k,
                    double kb);

/// Smallest eigenvalue of a symmetric 3x3 matrix (Jacobi sweeps).
double sym3_min_eig(const double a[9]);

// ------------------------------------------------------------------ solver

struct MasOracle {
  // padded topology + coarse space (mas_preconditioner.cu:99-478)
  int n = 0, padded = 0, level_num = 0;
  std::vector<int> real_to_padded, padded_to_real;
  std::vector<int> map;  // n * 4
  std::array<int, 4> cco_nums{};
  std::vector<int> mat_offsets, mat_nums;  // per level (level 0 = padded)
  std::vector<std::vector<double>> inv;    // dense 96x96 per matrix
  void build_topology(const Bsr3& A, const std::vector<int>& part_offsets);
  void factorize(const Bsr3& A, const std::vector<double>& diag);
  void apply(const double* r, double* z) const;
};

struct PcgResult {
  int iters = 0;
  int status = -1;
  double rr0 = 0, rr = 0;
};

/// mas_cg_solver.cu:264-382 in double.  criterion 0 stops on
/// rr <= rel^2 rr0 (the reference), 1 on rr <= rel^2 |b|^2.
PcgResult pcg_solve(const Bsr3& A, const std::vector<double>& diag,
                    const MasOracle& M, const double* b, double* x,
                    double rel_tol, double abs_tol, int max_iter,
                    int criterion = 0);

/// Clamp the singular values of a 3x2 matrix (column-major y[6]) to
/// [lo, hi] via a one-sided Jacobi SVD (cloth_admm_helper.cu:19-51).
void project_stretch(const double y[6], double lo, double hi, double out[6]);

// ------------------------------------------------------------------ collision

struct Box {
  float mn[3], mx[3];
};
inline bool box_overlap(const Box& a, const Box& b) {
  for (int d = 0; d < 3; ++d) {
    float lo = a.mn[d] > b.mn[d] ? a.mn[d] : b.mn[d];
    float hi = a.mx[d] < b.mx[d] ? a.mx[d] : b.mx[d];
    if (!(lo <= hi)) return false;
  }
  return true;
}

/// All (i, j) with box_overlap(A[i], B[j]) by sort-and-sweep on x (exact set
/// semantics of bbox.cuh:25-29 / oibvh.cuh traversal).  self=true returns
/// each unordered pair i < j of A once.
std::vector<std::pair<int, int>> overlap_pairs(const std::vector<Box>& A,
                                               const std::vector<Box>& B,
                                               bool self);

struct CcdConfig {
  double ms = 0.0;        // minimum separation
  double tolerance = 1e-6;
  long max_iter = 4096;
  double err[3] = {0, 0, 0};
  double max_time = 1.0;
  bool no_zero_toi = true;
  int stack_cap = 1024;   // pending-box capacity (overflow returns early)
  // 0: a root at t = 0 is searched again at ms / 2 only (ccd.cpp:150-180);
  // 1: then once more at zero separation.
  int zero_mode = 0;
};

/// Pending-box capacity of the world-level CCD; equals the HIP kernel's.
constexpr int kCcdStackCap = 64;

/// Earliest conservative TOI of a vertex-face (vf=true) or edge-edge query,
/// restating interval_root_finder.cpp:391-510 and ccd.cpp:195-257 in double.
/// Returns false for no hit; otherwise *toi in [0, max_time].  *zero (when
/// given) is set when the last search still had a root at t = 0 (no hit).
bool ccd_query(bool vf, const double p0[4][3], const double p1[4][3],
               const CcdConfig& cfg, double* toi, bool* zero = nullptr);

/// The same algorithm in single precision, as the reference CPU CCD runs it.
bool ccd_query_f32(bool vf, const float p0[4][3], const float p1[4][3],
                   float ms, float tolerance, long max_iter, const float err[3],
                   float max_time, float* toi);

/// Closest-point parameters (dcd.cu:321-394) in double.
bool pt_uv(const V3& x0, const V3& x1, const V3& x2, const V3& x3, double eps,
           double* u, double* v);
bool ee_uv(const V3& x0, const V3& x1, const V3& x2, const V3& x3, double eps,
           double* u, double* v);

/// Exact Signorini-Coulomb projection of one contact
/// (contact_constraints.cu:391-411).
void coulomb_solve(const double us[3], const double n[3], double mu,
                   double out[3]);

/// Numerical filter (interval_root_finder.cpp:512-522).
void ccd_filter(const double abs_max[3], bool vf, double err[3]);

}  // namespace oracle
