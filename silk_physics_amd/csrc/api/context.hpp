// Runtime context behind the C ABI: entity registry, lazily assembled device
// scene and the per-step orchestration (reference MainLoop::step,
// silk/src/backend/cuda/main_loop.cu:174-312, and ADMMSolver::solve,
// solver/admm_solver.cu:228-397).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "api/collision.hpp"
#include "api/devbuf.hpp"
#include "host/geometry.hpp"
#include "hip/kernels.hpp"
#include "hip/warm_start.hpp"
#include "silk_hip.h"

namespace silk::amd {

// Timing classes of silk_hip_get_timing.
enum : int {
  kTimeStep = 0,
  kTimePcg = 1,       // PCG solve kernel (persistent) per launch
  kTimeLocal = 2,     // local projection
  kTimeContact = 3,   // contact Gauss-Seidel
  kTimeBroad = 4,     // refit + traversal
  kTimeCcd = 5,       // CCD kernel
  kTimeResidual = 6,  // FP64 residual / RHS kernel
  kTimeHostContacts = 7,  // host wall time of contact assembly and colouring
  kTimeWarm = 8,      // learned warm start: features and network
};
constexpr int kTimeClasses = 9;

/// Restarts of a PCG solve under the true-residual criterion (the FP32
/// recurrence drifts from the FP64 residual by a small factor, so one or two
/// short restarts close the gap).
constexpr int kPcgRestarts = 3;

struct ClothEntity {
  uint32_t handle = 0;
  silk_hip_cloth_params cloth{};
  silk_hip_collision_params coll{};
  TriMesh mesh;           // caller order
  std::vector<int> pins;  // caller vertex indices
  bool has_pin_pos = false;
  std::vector<float> pin_curr;  // 3 * pins.size()
  // derived
  bool part_valid = false;
  Partition part;
  TriMesh pmesh;  // solver order
  bool geom_valid = false;
  ClothGeometry geom;
  // persistent state, caller order (valid when has_state)
  bool has_state = false;
  std::vector<float> x, v;
};

struct ObstacleEntity {
  uint32_t handle = 0;
  silk_hip_collision_params coll{};
  TriMesh mesh;
  bool has_pos = false;
  std::vector<float> curr, prev;
  bool is_static = true;
  bool is_static_twice = true;
};

/// Device scene: all cloths concatenated in solver order (one PCG system per
/// cloth) plus collision data.
struct Scene {
  int nv = 0, nf = 0, nsys = 0;
  float dt = 0.0f;
  std::vector<int> vert_off, face_off;  // per cloth (+1)
  int pinned_dof = 0;
  // ---- state and work vectors (3 * nv)
  DevBuf<float> x, v, prev_x, prev_v, outer_x, outer_v, inner_x, pin_lambda,
      dx;
  DevBuf<double> elastic_rhs;
  DevBuf<float> d;  // nv
  DevBuf<float> z[2], u[2];
  int aux_cur = 0;
  // ---- static cloth data
  DevBuf<int> F, vf_ptr, vf_list, pinned;
  DevBuf<float> fparam, fstiff, mass, damp, pin_target, contact_base_w;
  DevBuf<double> bend_rhs;
  std::vector<float> h_mass;  // nv, density applied
  std::vector<int> h_pinned;
  std::vector<double> h_contact_w;  // nv, Daviet base weight
  // ---- local-step partials
  DevBuf<double> admm_part;
  std::vector<double> h_admm_part;
  // ---- PCG / MAS
  DevBuf<int> blk_sys, blk_row0, blk_rows, sys_blk0, slice_off, slice_width,
      ell_col, bank_sys, bank_level, bank_member, cnode_ptr, cnode_list,
      cnode_level, row_cnode;
  DevBuf<float> ell_val, bank_static, bank_inv, bank_scale, r, p, q, z_pcg, zl0,
      cr, cz, cd;
  DevBuf<unsigned short> bank_inv_h;  // 16-bit bank inverses (mas_precision)
  DevBuf<float> x0_copy;    // 3*nv: start of the last solve (x0 of x0 + dx)
  DevBuf<double> r0d;       // 3*nv: its FP64 residual b - A x0
  DevBuf<double> dxc;       // 3*nv: FP64 committed correction (residual replacement)
  DevBuf<int> sys_done;     // per system: the true-residual test passed
  DevBuf<double> part3;     // 3 per row block: solve residual partials
  // ---- learned warm start (allocated when used)
  DevBuf<unsigned short> wfeat;  // 3*nv samples x 16 features
  DevBuf<float> wsig, wdelta, w_xprev, w_rprev;  // 3*nv
  DevBuf<double> w_part2;        // residual partials of the previous iterate
  DevBuf<double> w_r0dprev;      // 3*nv: its FP64 residual (true-residual test)
  DevBuf<double> w_gram;         // 30 per row block: projection sums
  DevBuf<float> w_coef;          // 6 per system: projection coefficients
  DevBuf<double> part, part2;
  std::vector<double> h_part2;
  DevBuf<PcgState> state;
  int nblk = 0, nbank = 0, ncnode = 0, n1node = 0;
  int64_t nnz_off = 0;              // off-diagonal scalar entries of K
  std::vector<int> h_bank_level;
  std::vector<int> h_blk0;  // sys_blk0 on host
  std::vector<PcgState> h_state;
  // ---- persistent PCG (one cooperative launch per solve)
  bool use_persistent = false;
  PersistDev pers;  // device pointers into the buffers below
  std::vector<std::pair<int, int>> ps_batches;  // (first workgroup, count) per launch
  // The batch lists of the one- and two-workgroups-per-CU variants (empty
  // where a system alone does not fit that way); ps_batches is one of them.
  std::vector<std::pair<int, int>> ps_variant[2];
  bool stepped = false;  // a step has run on this scene (autotune needs one)
  DevBuf<PcgSysDev> ps_sys;
  DevBuf<int> ps_slot_row, ps_sell_off, ps_sell_width, ps_sell_lcol, ps_halo_ptr,
      ps_halo_slot, ps_wg_n1_ptr, ps_wg_n1, ps_wg_c_ptr, ps_wg_c, ps_slot_cn,
      ps_cb_bank, ps_cb_member, ps_cb_rng, ps_abort, ps_slot_cnl, ps_halo_cnl;
  DevBuf<float> ps_sell_val;
  DevBuf<unsigned long long> ps_tz, ps_tn, ps_tp, ps_tc, ps_tcp, ps_tdx;
  DevBuf<unsigned> ps_bar;
  DevBuf<PcgState> ps_out;
  DevBuf<AdmmCtl> admm_ctl;
  AdmmCtl* h_admm_ctl = nullptr;  // pinned: AdmmCtl, then the abort flag
  // Executable graph of the device ADMM loop (admm_graph on): captured anew
  // for every outer iteration and updated in place while the captured
  // topology is unchanged; it holds this scene's buffer addresses, so it
  // lives and dies with the scene.
  hipGraphExec_t admm_exec = nullptr;
  Scene();
  ~Scene();
  // ---- collision
  std::unique_ptr<CollisionWorld> coll;
};

namespace cpu {
class CpuWorld;
}

class Context {
 public:
  explicit Context(int device);
  ~Context();

  int device_ = 0;
  // The CPU backend (device SILK_HIP_DEVICE_CPU): the scene is stepped by
  // cpu::CpuWorld, built from the entities at the first step after a change.
  bool cpu_ = false;
  std::unique_ptr<cpu::CpuWorld> cpu_world_;
  silk_hip_global cpu_global_{};  // the configuration cpu_world_ was built with
  void build_cpu_world();
  int cpu_step(silk_hip_step_stats* stats);
  std::string last_error_;
  silk_hip_global global_{};

  std::vector<ClothEntity> cloths_;
  std::vector<ObstacleEntity> obstacles_;
  uint64_t rng_state_ = 0x5117C0DE2024ULL;
  /// TOI (frame-time units) returned by the CCD of every outer iteration of
  /// the last step, in order.
  std::vector<float> toi_log_;
  std::vector<int> admm_log_;  // ADMM iterations of each outer iteration
  int last_admm_iters_ = 0;     // of the last admm_solve

  /// Learned warm start: network (FP32 master, Adam moments, BF16 copies),
  /// training scratch and the recorded samples.  Device buffers of the
  /// context's device; they outlive scene rebuilds.
  struct WarmNet {
    DevBuf<float> params, m, v, grad, grad_part;
    DevBuf<unsigned short> w1, w2, w3;
    DevBuf<double> loss_part, loss;
    bool loaded = false;
    int adam_t = 0;
    DevBuf<unsigned short> ds_feat;
    DevBuf<float> ds_target;
    DevBuf<unsigned long long> ds_count;
    int64_t ds_cap = 0;
    float ds_rate = 0.0f;
    uint32_t ds_salt = 0;
    MlpDev dev() const;
  };
  WarmNet warm_;
  bool has_warm_net() const { return warm_.loaded; }
  bool recording() const { return warm_.ds_cap > 0; }
  /// Allocates the network buffers (device must be ready).
  void warm_alloc();
  void warm_set(const float* params);
  void warm_get(float* params);
  void warm_record(int64_t capacity, float rate);
  int64_t warm_samples();
  void warm_train(int steps, int batch, float lr, uint64_t seed, float* loss);
  void warm_debug(int64_t n, const float* feat, const float* target, float* y,
                  float* grad, double* loss);

  bool dev_ready_ = false;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev_[2] = {nullptr, nullptr};
  bool scene_dirty_ = true;
  std::unique_ptr<Scene> scene_;
  double timing_ms_[kTimeClasses] = {};
  int64_t timing_launch_[kTimeClasses] = {};
  // Per-kernel event timing (off unless enabled): pairs of events recorded
  // around launches on stream_ of the classes whose bit is set, resolved
  // after the next synchronisation.
  uint32_t timing_mask_ = 0;
  bool timed_open_ = false;  // the last timed_begin recorded its event
  bool timing(int cls) const { return ((timing_mask_ >> cls) & 1u) != 0; }
  // A class of the ADMM loop is timed (events between its launches).
  bool timing_admm() const {
    return timing(kTimePcg) || timing(kTimeLocal) || timing(kTimeContact) ||
           timing(kTimeResidual) || timing(kTimeWarm);
  }
  struct EvPair {
    hipEvent_t a = nullptr, b = nullptr;
  };
  std::vector<EvPair> ev_pool_;
  std::vector<int> ev_cls_;  // class of each recorded pair
  void timed_begin(int cls);
  void timed_end();
  void timed_collect();  // stream must be synchronised

  uint32_t new_handle();
  ClothEntity* find_cloth(uint32_t h);
  ObstacleEntity* find_obstacle(uint32_t h);

  /// Download cloth states from the device scene, then drop it.
  void invalidate_scene();
  void ensure_device();
  void ensure_cloth_derived(ClothEntity& c);
  void build_scene();
  int step(silk_hip_step_stats* stats);
  int pcg_solve(float rel_tol, bool factor, const RhsArgs& ra,
                silk_hip_step_stats* st, bool restart = false);
  /// One persistent PCG solve: a launch per batch of systems.
  void launch_persistent(PersistDev a, float rel_tol);
  /// True-residual criterion on the host-driven ADMM path: tests the FP64
  /// residual of the last solve's iterate and restarts the solve from the
  /// committed x while it is above rel_tol |b| (at most kPcgRestarts times).
  void verify_host(float rel_tol, const RhsArgs& ra, silk_hip_step_stats* stats);
  /// Multi-launch PCG iterations (used when the persistent kernel cannot
  /// keep its grid resident); leaves per-system states in scene h_state.
  void pcg_iterate(float rel_tol, const RhsArgs& ra);
  int finish_pcg(int iters, double rel, const RhsArgs& ra,
                 silk_hip_step_stats* st, bool restart = false);
  bool admm_solve(silk_hip_step_stats* st);
  bool admm_solve_device(silk_hip_step_stats* st);
  // Enqueues iterations [it0, it1) of one outer iteration's device ADMM loop
  // on stream_ (kernels only), with the loop's set-up when it0 == 0 and its
  // final pin enforcement when it1 is the last.
  void admm_enqueue(bool projected, bool learned, bool record, int it0, int it1);
  // Reads the counters of the last enqueued device ADMM loop (synchronises
  // the stream; a no-op when none is pending); false when it diverged.
  bool admm_device_finish(silk_hip_step_stats* st);
  bool admm_pending_ = false;
  // The persistent PCG serves the scene's solves: it fits the device, and
  // not BF16 banks in the two-workgroups-per-CU variant, whose solves
  // diverged on the 256x256 sheet (cause not found; those solves take the
  // multi-launch PCG instead).
  bool persistent_on() const {
    return scene_ && scene_->use_persistent &&
           !(scene_->pers.dense && global_.mas_precision == SILK_HIP_MAS_BF16);
  }
  // Ends the stream capture of the ADMM loop, updates (or instantiates) the
  // scene's executable graph from it and launches that on stream_.
  void launch_admm_graph(silk_hip_step_stats* st);
  void get_positions(const ClothEntity& c, float* out) const;
  /// silk_hip_autotune: times the broadphase split depths and the PCG launch
  /// variants on the current scene and keeps the fastest (report layout in
  /// silk_hip.h).
  int autotune(int32_t* report);
  /// silk_hip_debug_pcg: one solve of K + diag(d) x = b from x = 0.
  /// launch_ms, when given, receives the persistent launch's duration.
  int debug_pcg(const float* d, const float* b, float rel_tol, float* x,
                float* launch_ms = nullptr);
  /// Push one cloth's pin targets to the device scene (if built).
  void update_pin_targets(const ClothEntity& c);

  // global config with defaults resolved
  int pcg_max_iter() const { return global_.pcg_max_iter > 0 ? global_.pcg_max_iter : 1000; }
  int contact_sweeps() const { return global_.contact_sweeps > 0 ? global_.contact_sweeps : 10; }
  int ccd_max_iter() const { return global_.ccd_max_iter > 0 ? global_.ccd_max_iter : 4096; }
  float ccd_tolerance() const { return global_.ccd_tolerance > 0 ? global_.ccd_tolerance : 1e-6f; }
  int ccd_zero_mode() const { return global_.ccd_zero_toi; }
  float ccd_ms_scale() const { return global_.ccd_min_separation_scale > 0 ? global_.ccd_min_separation_scale : 0.2f; }
  float dcd_scale() const { return global_.dcd_activation_scale > 0 ? global_.dcd_activation_scale : 1.0f; }

  ClothDev cloth_dev();
  PcgDev pcg_dev();
};

/// silk_hip_ccd_min_noalloc (hip/ccd_noalloc.hip).
int ccd_min_noalloc(Context& c, int64_t n, const int32_t* is_vf, const double* p0,
                    const double* p1, const double* err, double ms, double tolerance,
                    int32_t max_iter, double* min_toi);

}  // namespace silk::amd
