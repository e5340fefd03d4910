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
  const double* bend_rhs = nullptr;  // 3*nv, kb * CWC * V
  const int* pinned = nullptr;       // nv: 1 if pinned
  const float* pin_target = nullptr; // 3*nv
};

/// prev_x = x; prev_v = damp*v; outer_* = prev_*  (main_loop.cu:58-110).
void launch_step_init(hipStream_t s, const ClothDev& c, const float* x,
                      const float* v, float* prev_x, float* prev_v,
                      float* outer_x, float* outer_v);

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
  double* elastic_rhs = nullptr;  // 3*nv, next main rhs elastic part
  float* pin_lambda = nullptr;   // 3*nv
  // contact terms accumulated into the dual residual (may be null)
  const float* contact_dual = nullptr;   // 3*nv: w (p - p_prev)
  const float* contact_eta = nullptr;    // 3*nv
  const float* contact_w = nullptr;      // nv: active weight
  const float* contact_p = nullptr;      // 3*nv
  double* partials = nullptr;            // kAdmmPartials * grid
  // When non-null and *skip != 0 the launch only carries z and u over
  // (z_out = z_in, u_out = u_in): the ADMM loop has already stopped.
  const int* skip = nullptr;
};

/// Device-side control of one ADMM solve (admm_solver.cu:209-330): the
/// stopping test runs on the device after every local step, and every
/// kernel of a later iteration becomes a no-op once `done` is set, so the
/// host enqueues all iterations without waiting on any of them.
struct AdmmCtl {
  int done = 0;       // stopping test passed, or diverged
  int diverged = 0;   // non-finite state seen
  int admm_iters = 0;
  int pcg_iters = 0;
  int pcg_solves = 0;
  int pcg_max_solve = 0;  // largest iteration count of one solve (atomicMax)
  int pcg_stalled = 0;    // solves that ended at max_iter
  int warm_kept = 0;      // learned warm starts kept by the judge
  int warm_solves = 0;    // learned warm starts evaluated
  int pcg_restarts = 0;   // true-residual criterion: restarts of a system
  int verify_done = 0;    // true-residual criterion: every system's iterate
                          // meets the tolerance (or the ADMM loop stopped)
  int pcg_unverified = 0; // solves still above the tolerance after the last
                          // restart
  int verify_arrive = 0;  // row blocks done in the verification kernel
  float rel_tol = 0;  // linear tolerance of the next x-update
  float last_rel = 0; // sqrt(rr / rr0) of the last PCG solve
  double init_primal = 0, init_dual = 0, ratio = 1;
  // Largest |b - A x|^2 / |b|^2 over the solves (bits of a non-negative
  // double, so that an integer atomicMax orders them): the PCG iterate
  // x0 + dx and the stored FP32 x.
  unsigned long long max_true_bits = 0, max_commit_bits = 0;
};

struct AdmmCheckArgs {
  int grid = 0;        // local-step workgroups (partials)
  int n3 = 0;          // 3 * cloth vertices
  int primal_dof = 0;  // 6 F + pinned dof + active contact dof
  int it = 0;          // ADMM iteration just finished
  float abs_tol = 0, rel_tol = 0;
  float lin_init = 0, lin_min = 0, lin_max = 0;
};

/// ctl = {not done, rel_tol = clamp(lin_init)} before an ADMM solve.
void launch_admm_init(hipStream_t s, AdmmCtl* ctl, float lin_init,
                      float lin_min, float lin_max);
/// Stopping test of iteration a.it from the local-step partials
/// (admm_solver.cu:296-330): updates ratio, rel_tol and done.
void launch_admm_check(hipStream_t s, const double* partials,
                       const AdmmCheckArgs& a, AdmmCtl* ctl);

/// ADMM aux-variable and multiplier update for cloth, pins and the contact
/// primal term (cloth_admm_helper.cu:80-262, equality_constraints.cu:90-154,
/// contact_constraints.cu:316-364) in one vertex-centric pass.
void launch_local_step(hipStream_t s, const ClothDev& c,
                       const LocalStepArgs& a);
int local_step_grid(int nv);

/// Pins x = t, contact vertices x = p (equality_constraints.cu:156-167,
/// contact_constraints.cu:367-374).
void launch_enforce(hipStream_t s, const ClothDev& c, const float* contact_w,
                    const float* contact_p, float* x);

// ------------------------------------------------------------------- PCG

/// Per-system PCG state (one system per cloth).  Double-buffered by iteration
/// parity: iteration k reads state[k&1] and writes state[(k+1)&1].
struct PcgState {
  double rz, rr, rr0, bb;
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
  float* bank_inv = nullptr;           // 1024 per bank (fp32)
  int ncnode = 0;                      // coarse nodes, all levels and systems
  const int* cnode_ptr = nullptr;      // ncnode+1 : members (rows for level 1,
  const int* cnode_list = nullptr;     //   level-1 nodes for deeper levels)
  const int* cnode_level = nullptr;    // ncnode
  const int* row_cnode = nullptr;      // 4 per row: coarse node per level, -1
  int n1node = 0;                      // level-1 nodes are ids [0, n1node)
  // vectors (3 per row)
  float* x = nullptr;   // warm start x0 (updated by launch_pcg_commit)
  float* dx = nullptr;  // accumulated correction: solution = x0 + dx
  float* r = nullptr;
  float* p = nullptr;
  float* q = nullptr;
  float* z = nullptr;
  float* zl0 = nullptr;   // level-0 bank output (3 per row)
  float* cr = nullptr;    // coarse r (3 per coarse node)
  float* cz = nullptr;    // coarse z (3 per coarse node)
  float* d = nullptr;        // diag (per row), written by launch_pcg_init
  double* part = nullptr;    // 2 doubles per block
  double* part2 = nullptr;   // 2 doubles per block (residual kernel)
  PcgState* state = nullptr; // 2*nsys
  float abs_tol = 0.0f;
  int max_iter = 1000;
  int criterion = 0;         // 0: rr <= rel^2 rr0; 1: rr <= rel^2 |b|^2;
                             // 2: as 1, then the FP64 residual of the iterate
                             // is checked and the solve restarted while it
                             // is above rel |b|
  int mas_half = 0;          // bank inverses read from bank_inv_h (0 = fp32,
                             // 1 = fp16, 2 = bf16), written by the factor
  unsigned short* bank_inv_h = nullptr;  // 1024 per bank
  float* bank_scale = nullptr;           // per bank (fp16: entries / scale)
};

/// Factorize: bank matrices = static + aggregated diag, inverted in place
/// (mas_preconditioner.cu:489-598,730-859).
void launch_mas_factor(hipStream_t s, const PcgDev& p, float* cd_scratch,
                       const int* skip = nullptr);

/// Right-hand side of the x-update (admm_solver.cu:272-292,
/// cloth_admm_helper.cu:264-332, equality_constraints.cu:107-120,
/// contact_constraints.cu:299-314), consumed directly by the residual kernel.
struct RhsArgs {
  ClothDev c;
  const float* prev_x = nullptr;       // frame-start state (inertia term)
  const float* prev_v = nullptr;
  const double* elastic_rhs = nullptr; // from the local step
  const float* pin_lambda = nullptr;
  float pin_rho = 1e4f;
  const float* cw = nullptr;           // active contact weight (may be null)
  const float* cp = nullptr;
  const float* ceta = nullptr;
  const int* skip = nullptr;           // no-op when *skip != 0
  float* x_copy = nullptr;             // optional: the residual kernel copies
                                       //   x here (the solve's start x0)
  double* r0d = nullptr;               // optional: the FP64 residual per row
};

/// PCG init, part 1 (mas_cg_solver.cu:273-277): d and r0 = b - (K + D) x0
/// assembled and evaluated in FP64 from FP32 inputs.  Must precede
/// launch_mas_factor, which reads d.
void launch_pcg_residual(hipStream_t s, const PcgDev& p, const RhsArgs& ra);

/// PCG init, part 2 (mas_cg_solver.cu:279-296): z = M r0, p = z, rz0/rr0,
/// dx = 0; with criterion 1 a system whose r0 already meets rel_tol * |b|
/// is done at iteration 0.
void launch_pcg_start(hipStream_t s, const PcgDev& p, float rel_tol);

/// FP64 true residual b - (K + D) x of the committed solution: per-block
/// partials (|r|^2, |b|^2) in p.part2; r written to p.q.
void launch_true_residual(hipStream_t s, const PcgDev& p, const RhsArgs& ra);

/// FP64 residuals of a finished solve (b from ra): the iterate x0 + dx and
/// the stored xc, relative to |b|, per system; their squares are folded into
/// ctl->max_true_bits (when fold_iterate) / max_commit_bits.  part3: 3
/// doubles per row block.
/// No-op when *ra.skip != 0.
void launch_solve_residuals(hipStream_t s, const PcgDev& p, const RhsArgs& ra,
                            const float* x0, const float* dx, const float* xc,
                            double* part3, AdmmCtl* ctl, int fold_iterate);

/// True-residual stopping test of a finished solve: per system, the FP64
/// residual of the iterate x0 + dx (never rounded to FP32) against
/// max(abs_tol, rel * |b|), rel = ctl->rel_tol (or `rel_tol` when
/// ctl->rel_tol is 0).  Writes sys_done[sys] and ctl->verify_done (1 when
/// every system passes or ctl->done is set).  Systems that pass, and with
/// `final` set every system, fold their ratio into ctl->max_true_bits;
/// with `final` set, failing systems are counted in ctl->pcg_unverified.
/// part3: 2 doubles per row block.  No-op when ctl->done is set (then only
/// verify_done is written).
/// With `first` unset, systems whose sys_done is already set keep it and are
/// not evaluated again (their dx was not written by the restart).
void launch_iterate_verify(hipStream_t s, const PcgDev& p, const RhsArgs& ra,
                           const float* x0, const float* dx, double* part3,
                           int* sys_done, AdmmCtl* ctl, float rel_tol, int first,
                           int final);

/// dst = src (n floats); no-op when *skip != 0.
void launch_copy_unless(hipStream_t s, float* dst, const float* src, int64_t n,
                        const int* skip);

/// p[i] = v for n 32-bit or 64-bit words.  Used instead of hipMemsetAsync /
/// device-to-device hipMemcpyAsync inside the step: the runtime's blit
/// kernels leave the device idle for several microseconds around each one,
/// a kernel of the library's follows the previous one back to back.
void launch_fill_u32(hipStream_t s, void* p, int64_t n, uint32_t v);
void launch_fill_u64(hipStream_t s, void* p, int64_t n, unsigned long long v);

/// x0 += dx.
void launch_pcg_commit(hipStream_t s, const PcgDev& p);

/// PCG iterations k .. k+count-1 (mas_cg_solver.cu:321-357 body).
void launch_pcg_iters(hipStream_t s, const PcgDev& p, int k, int count,
                      float rel_tol);

/// Apply the MAS preconditioner once to p.r -> p.z (tests).
void launch_mas_apply(hipStream_t s, const PcgDev& p);

// ------------------------------------------------- persistent PCG solve
//
// One launch runs a whole PCG solve (all iterations) for every system.  Each
// system owns a group of workgroups; a workgroup owns up to eight 32-slot
// level-0 MAS banks, keeps the rows of those banks (r, dx, z, p, q) in
// registers and the inverses of its banks and of all coarse banks of its
// system in LDS for the whole solve.  Only z, p, the per-node sums of q and
// the dot-product partials travel through memory, written and read with
// device-coherent (agent-scope) atomics.  Workgroups of a group meet at two
// grid barriers per iteration (p.Ap; r.z and r.r); the level-1 residual sums
// that the coarse levels need are kept current in every workgroup by the
// recurrence cr -= alpha * (sum of q over the node).

constexpr int kPcgThreads = 256;       // threads per workgroup (8 banks x 32)
constexpr int kPcgBarrierStride = 64;  // unsigned words per system barrier
constexpr int kPcgLdsBytes = 150 * 1024;  // dynamic LDS budget per workgroup
constexpr int kPcgNbr = 32;        // neighbours kept in registers per row
constexpr int kPcgNbrChunk = 8;    // slot-ELL widths are multiples of this

struct PcgSysDev {
  int slot0 = 0, nslot = 0;  // padded slots (32 per level-0 bank)
  int bank0 = 0;             // global bank id of the first level-0 bank
  int cb0 = 0, ncb = 0;      // coarse banks in the persistent coarse list
  int cn0 = 0, ncn = 0;      // coarse nodes (system-local ids cn0 + local)
  int n1 = 0;                // level-1 nodes (local ids [0, n1))
  int lvl[5] = {0, 0, 0, 0, 0};  // level l nodes: local ids [lvl[l-1], lvl[l])
  int wg0 = 0, nwg = 0;      // slot workgroups of the group
  int ncw = 0;               // coarse workgroups, after the slot ones: coarse
                             // bank cb is solved by coarse workgroup cb % ncw
};

struct PersistDev {
  int nsys = 0;
  int nslot_total = 0;              // slots of all systems
  const PcgSysDev* sys = nullptr;
  const int* slot_row = nullptr;    // global row of a slot, -1 for padding
  const int* sell_off = nullptr;    // slot-ELL: 64-slot slices
  const int* sell_width = nullptr;
  const int* sell_lcol = nullptr;   // neighbour as an index into the owning
                                    // workgroup's slots [0, 256) then halo
  const float* sell_val = nullptr;
  const int* halo_ptr = nullptr;    // per workgroup: its halo slots
  const int* halo_slot = nullptr;   //   (other workgroups' slots it reads)
  const int* wg_n1_ptr = nullptr;   // per coarse workgroup: the level-1 nodes
  const int* wg_n1 = nullptr;       //   (global ids) its banks read, member by
                                    //   member
  const int* wg_c_ptr = nullptr;    // per workgroup: coarse nodes whose
  const int* wg_c = nullptr;        //   corrections its slots and halo add
  const int* slot_cn = nullptr;     // 4 per slot: coarse node (cn0-relative) per level
  const int* slot_cnl = nullptr;    // 4 per slot: the same nodes as positions in the
                                    //   owner workgroup's wg_c list, -1 if absent
  const int* halo_cnl = nullptr;    // 4 per halo entry: positions in the reading
                                    //   workgroup's wg_c list, -1 if absent
  const int* cb_bank = nullptr;     // per coarse bank: global bank id (bank_inv)
  const int* cb_member = nullptr;   // 32 per coarse bank: node (cn0-relative), -1
  const int* cb_rng = nullptr;      // 2 per coarse bank member: its nodes'
                                    //   positions [begin, end) in the list
  const float* bank_inv = nullptr;  // 1024 per bank
  const unsigned short* bank_inv_h = nullptr;  // 1024 per bank, half/bf16
  const float* bank_scale = nullptr;  // per bank
  int mas_half = 0;                 // 0 fp32 (bank_inv), 1 fp16, 2 bf16
  const float* d = nullptr;         // diag per row
  const float* r0 = nullptr;        // FP32 initial residual per row
  float* x = nullptr;               // x += dx at the end (rows)
  // Epoch-tagged exchange words (epoch << 32 | payload bits).
  unsigned long long* tz = nullptr;  // 3 per slot: z
  unsigned long long* tn = nullptr;  // 3 per coarse node: level-1 sums
  unsigned long long* tp = nullptr;  // kPcgPartWords per workgroup: partials
  unsigned long long* tc = nullptr;  // 3 per coarse node: coarse correction
  unsigned long long* tcp = nullptr; // 2 per coarse bank: its r.z_coarse
  unsigned* bar = nullptr;          // kPcgBarrierStride per system: epoch
  int* abort_flag = nullptr;
  PcgState* out = nullptr;          // per system
  AdmmCtl* ctl = nullptr;           // optional: skip flag, tolerance, counters
  const int* sys_blk0 = nullptr;    // per system: first row block (nsys+1)
  const double* part2 = nullptr;    // per row block: (|r0|^2, |b|^2) from the
                                    //   residual kernel
  float* dx_out = nullptr;          // optional: the solve's dx (rows)
  const int* sys_done = nullptr;    // optional, per system: nonzero = already
                                    //   converged, the launch leaves it alone
  // True-residual test inside the launch (criterion 2): when r0d is set, a
  // solve that meets the recurrence test computes the FP64 residual
  // r0d - A dx of its iterate (dx of the halo exchanged through tdx) and,
  // above max(abs^2, rel^2 |b|^2), replaces r by it and continues, at most
  // max_replace times.  A replacement moves dx into the FP64 sum dxc and
  // the FP64 residual into r0d, so the search continues from dx = 0 and the
  // iterate x0 + dxc + dx keeps FP64 precision in its committed part.
  double* r0d = nullptr;            // FP64 b - A x0 per row (residual kernel)
  double* dxc = nullptr;            // FP64 committed correction per row
  unsigned long long* tdx = nullptr;  // 3 per slot: dx, epoch-tagged
  int max_replace = 0;
  float abs_tol = 0.0f;
  int max_iter = 1000;
  int criterion = 0;                // 0: rr <= rel^2 rr0; 1 and 2: rr <= rel^2 |b|^2
  int restart = 0;                  // 1: a restart of unconverged systems
                                    //   (counted in ctl->pcg_restarts)
  int need_max = 0;     // max level-1 nodes listed for one workgroup (LDS)
  int ncn_max = 0;      // max coarse nodes of a system
  int nc_max = 0;       // max coarse nodes one workgroup's slots add (LDS)
  int ncb_max = 0;      // max coarse banks of a system (LDS)
  int nhalo_max = 0;    // max halo slots of a workgroup (LDS)
  int nwg_max = 0;      // max workgroups of a system (LDS)
  int grid = 0;         // workgroups of one launch
  int wg_base = 0;      // first workgroup (over all systems) of the launch
  int dense = 0;        // 1: the two-workgroups-per-CU variant
};

/// Tagged words of an exchange one thread of the persistent kernel keeps the
/// addresses of in registers; the words beyond kPcgThreads * kPcgXchWords are
/// mapped when collected, kPcgXchBatch per thread at a time.
constexpr int kPcgXchWords = 8;
constexpr int kPcgXchBatch = 4;
/// Wavefronts per persistent-PCG workgroup; each publishes its own partial
/// sums (two words per double): per workgroup, p.q at [0, 8), r.z0 at
/// [8, 16), r.r at [16, 24) of kPcgPartWords words.
constexpr int kPcgWaves = kPcgThreads / 64;
constexpr int kPcgPartWords = 32;
/// Halo slots one thread completes (a workgroup's halo has at most
/// kPcgThreads * kPcgHaloPerThread slots).
constexpr int kPcgHaloPerThread = 3;

/// Dynamic LDS bytes of the persistent kernel for the sizes in `a`.
size_t pcg_persistent_lds(const PersistDev& a);

/// True when the device can keep `grid` workgroups of the persistent kernel
/// resident at once with the given dynamic LDS (dense: the variant that
/// shares a CU between two workgroups).
bool pcg_persistent_fits(int grid, size_t lds_bytes, int dense);

/// One whole PCG solve (mas_cg_solver.cu:193-311) per system, x += solution.
hipError_t launch_pcg_persistent(hipStream_t s, const PersistDev& a,
                                 float rel_tol);

}  // namespace silk::amd
