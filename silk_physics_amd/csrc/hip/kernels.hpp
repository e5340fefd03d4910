// Launch wrappers for the HIP kernels of the PD step.  Device pointers only;
// sizes are validated by the caller (context.cpp) before every launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace silk::amd {

constexpr int kBlock = 256;     // threads per workgroup for row kernels
constexpr int kSlice = 64;      // rows per sliced-ELL slice (one wavefront)
constexpr int kBankSlots = 32;  // MAS bank size

// ---------------------------------------------------------------- PD / ADMM

struct ClothDev {
  int nv = 0;  // vertices (all cloths, solver order)
  int nf = 0;  // faces
  float h = 0.0f;
  float acc[3] = {0, 0, 0};
  const int* F = nullptr;            // 3*nf
  const float* fparam = nullptr;     // 4*nf: ia, ib, ic, sqrt(area)
  const float* fstiff = nullptr;     // 2*nf: elastic stiffness k, penalty rho
  const int* vf_ptr = nullptr;       // nv+1
  const int* vf_list = nullptr;      // face*4 + corner
  const float* mass = nullptr;       // nv (density applied)
  const float* damp = nullptr;       // nv: 1 - damping
  const float* bend_rhs = nullptr;   // 3*nv
  const int* pinned = nullptr;       // nv: 1 if pinned
  const float* pin_target = nullptr; // 3*nv
};

/// prev_x = x; prev_v = damp*v; outer_* = prev_*; minert = m(prev_x/h^2 +
/// prev_v/h + a)  (main_loop.cu:58-110, admm_solver.cu:141-151).
void launch_step_init(hipStream_t s, const ClothDev& c, const float* x,
                      const float* v, float* prev_x, float* prev_v,
                      float* outer_x, float* outer_v, float* minert);

/// inner = (outer + h v) + h^2 a  (main_loop.cu:112-132).
void launch_predict(hipStream_t s, const ClothDev& c, const float* outer_x,
                    const float* outer_v, float* inner);

/// out = (b_end - a_start) / h  (main_loop.cu:134-143).
void launch_velocity(hipStream_t s, int n3, float h, const float* a_start,
                     const float* b_end, float* out);

/// out = w*a + (1-w)*b  (main_loop.cu:145-154).
void launch_mix(hipStream_t s, int n3, float w, const float* a, const float* b,
                float* out);

/// Per-workgroup partials written by launch_local_step (8 doubles each):
/// 0 ||W(Sx-z)||^2 + pin/contact primal, 1 ||W Sx||^2 + ||x_pin||^2,
/// 2 ||W z||^2 + ||t_pin||^2, 3 ||dual||^2, 4 ||dual_curr||^2,
/// 5 ||dual_prev||^2, 6 non-finite count, 7 unused.
constexpr int kAdmmPartials = 8;

struct LocalStepArgs {
  float max_mul = 1e8f;
  float pin_rho = 1e4f;
  const float* x = nullptr;
  const float* z_in = nullptr;
  const float* u_in = nullptr;
  float* z_out = nullptr;
  float* u_out = nullptr;
  float* elastic_rhs = nullptr;  // 3*nv, next main rhs elastic part
  float* pin_lambda = nullptr;   // 3*nv
  // contact terms accumulated into the dual residual (may be null)
  const float* contact_dual = nullptr;   // 3*nv: w (p - p_prev)
  const float* contact_eta = nullptr;    // 3*nv
  const float* contact_w = nullptr;      // nv: active weight
  const float* contact_p = nullptr;      // 3*nv
  double* partials = nullptr;            // kAdmmPartials * grid
};

/// ADMM aux-variable and multiplier update for cloth, pins and the contact
/// primal term (cloth_admm_helper.cu:80-262, equality_constraints.cu:90-154,
/// contact_constraints.cu:316-364) in one vertex-centric pass.
void launch_local_step(hipStream_t s, const ClothDev& c,
                       const LocalStepArgs& a);
int local_step_grid(int nv);

/// Main-update system: diag d (per vertex) and rhs b (3 per vertex)
/// (admm_solver.cu:272-292, cloth_admm_helper.cu:264-332,
/// equality_constraints.cu:107-120, contact_constraints.cu:299-314).
void launch_assemble(hipStream_t s, const ClothDev& c, const float* minert,
                     const float* elastic_rhs, const float* pin_lambda,
                     const float* contact_w, const float* contact_p,
                     const float* contact_eta, float* d, float* b);

/// Pins x = t, contact vertices x = p (equality_constraints.cu:156-167,
/// contact_constraints.cu:367-374).
void launch_enforce(hipStream_t s, const ClothDev& c, const float* contact_w,
                    const float* contact_p, float* x);

// ------------------------------------------------------------------- PCG

/// Per-system PCG state (one system per cloth).  Double-buffered by iteration
/// parity: iteration k reads state[k&1] and writes state[(k+1)&1].
struct PcgState {
  double rz, rr, rr0, pAp;
  int iter;
  int done;     // 1 once converged / stopped
  int status;   // 0 abs, 1 rel, 2 max iter, 3 invalid initial residual, -1 running
  int pad;
};

struct PcgDev {
  int nsys = 0;
  int nv = 0;
  // row blocks (256 rows, never straddle systems)
  int nblk = 0;
  const int* blk_sys = nullptr;
  const int* blk_row0 = nullptr;
  const int* blk_rows = nullptr;
  const int* sys_blk0 = nullptr;  // nsys+1
  // sliced ELL of the vertex-scalar operator (slices of 64 rows inside blocks)
  const int* slice_off = nullptr;   // per slice: offset into ell arrays
  const int* slice_width = nullptr; // per slice
  const int* ell_col = nullptr;
  const float* ell_val = nullptr;
  // MAS
  int nbank = 0;
  const int* bank_sys = nullptr;
  const int* bank_level = nullptr;     // 0 = fine, >0 coarse
  const int* bank_member = nullptr;    // 32 per bank: row (L0) or node (coarse), -1 pad
  const float* bank_static = nullptr;  // 1024 per bank
  float* bank_inv = nullptr;           // 1024 per bank (fp32) or fp16 view
  int ncnode = 0;                      // coarse nodes, all levels and systems
  const int* cnode_ptr = nullptr;      // ncnode+1 : members (rows for level 1,
  const int* cnode_list = nullptr;     //   level-1 nodes for deeper levels)
  const int* cnode_level = nullptr;    // ncnode
  const int* row_cnode = nullptr;      // 4 per row: coarse node per level, -1
  int n1node = 0;                      // level-1 nodes are ids [0, n1node)
  // vectors (3 per row)
  float* x = nullptr;
  float* r = nullptr;
  float* p = nullptr;
  float* q = nullptr;
  float* z = nullptr;
  float* zl0 = nullptr;   // level-0 bank output (3 per row)
  float* cr = nullptr;    // coarse r (3 per coarse node)
  float* cz = nullptr;    // coarse z (3 per coarse node)
  const float* d = nullptr;  // diag (per row)
  const float* b = nullptr;  // rhs
  double* part = nullptr;    // 2 doubles per block
  PcgState* state = nullptr; // 2*nsys
  const float* rel_tol = nullptr;  // per system? (one value for all)
  float abs_tol = 0.0f;
  int max_iter = 1000;
};

/// Factorize: bank matrices = static + aggregated diag, inverted in place
/// (mas_preconditioner.cu:489-598,730-859).
void launch_mas_factor(hipStream_t s, const PcgDev& p, float* cd_scratch);

/// PCG init: r = b - A x, z = M r, p = z, rz0/rr0  (mas_cg_solver.cu:264-296).
void launch_pcg_init(hipStream_t s, const PcgDev& p, float rel_tol);

/// PCG iterations k .. k+count-1 (mas_cg_solver.cu:321-357 body).
void launch_pcg_iters(hipStream_t s, const PcgDev& p, int k, int count,
                      float rel_tol);

/// Apply the MAS preconditioner once to p.r -> p.z (tests).
void launch_mas_apply(hipStream_t s, const PcgDev& p);

}  // namespace silk::amd
