/* C ABI of the MI355X projective-dynamics step.
 *
 * This is the seam between the C++20 silk::World front end and the HIP device
 * library.  It replaces the reference's internal backend interface
 * (silk/src/common/backend.hpp:9-45, implemented by CudaBackend in
 * silk/src/backend/cuda/cuda_backend.cu:64-352): every World call maps to one
 * function here.  Signatures use plain pointers, sizes and POD structs only.
 *
 * Conventions
 *  - Every function returns an int status: SILK_HIP_OK (0) or an error code
 *    whose value equals the silk::ErrorCode enumerator + 1.
 *  - Caller buffers are read (or written) during the call only and are never
 *    retained.
 *  - Calls are synchronous: the device stream is drained before returning.
 *  - One context is driven from one thread at a time.
 *  - Device and runtime faults never unwind across the ABI: they return
 *    SILK_HIP_ERR_UNKNOWN and silk_hip_last_error() describes them.
 *  - Handles are non-zero, unique within a context (reference:
 *    silk/src/ecs/registry.hpp:70-85).
 *  - Device work is deferred to the first silk_hip_step after a change, as
 *    in the reference (silk/src/backend/cuda/main_loop.cu:58-110).
 */
#ifndef SILK_HIP_H_
#define SILK_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SILK_HIP_OK = 0,
  SILK_HIP_ERR_UNKNOWN = 1,
  SILK_HIP_ERR_NOT_INITIALIZED = 2,
  SILK_HIP_ERR_INVALID_CONFIG = 3,
  SILK_HIP_ERR_TOO_MANY_BODY = 4,
  SILK_HIP_ERR_INVALID_HANDLE = 5,
  SILK_HIP_ERR_INVALID_MESH = 6,
  SILK_HIP_ERR_INVALID_PIN = 7,
  SILK_HIP_ERR_INVALID_POSITION = 8,
  SILK_HIP_ERR_CHOLESKY_FAIL = 9,
  SILK_HIP_ERR_NEED_INIT_SOLVER = 10,
  SILK_HIP_ERR_ITERATIVE_SOLVE_FAIL = 11,
  SILK_HIP_ERR_NO_DEVICE = 12
};

typedef struct silk_hip_ctx silk_hip_ctx;

/* silk_hip_create's device number of the CPU backend. */
#define SILK_HIP_DEVICE_CPU (-1)

/* silk::ClothConfig (silk/include/silk/silk.hpp:37-47). */
typedef struct {
  float elastic_stiffness;
  float bending_stiffness;
  float density;
  float damping;
} silk_hip_cloth_params;

/* silk::CollisionConfig (silk/include/silk/silk.hpp:28-34). */
typedef struct {
  int32_t is_collision_on;
  int32_t is_self_collision_on;
  int32_t group;
  float restitution;
  float friction;
} silk_hip_collision_params;

/* silk::GlobalConfig (silk/include/silk/silk.hpp:50-77), plus solver knobs
 * that the reference keeps as hidden constants (main_loop.cuh:21-26,
 * contact_constraints.cuh:89, mas_cg_solver.cuh:22).  Fields after
 * admm_rel_tol are extensions: zero selects the reference value. */
typedef struct {
  float acceleration[3];
  int32_t max_outer_iteration;
  int32_t max_inner_iteration;
  float dt;
  float linear_solver_abs_tol;
  float linear_solver_rel_tol_min;
  float linear_solver_rel_tol_max;
  float initial_linear_rel_tol;
  float admm_abs_tol;
  float admm_rel_tol;
  /* Extensions (0 = reference default). */
  int32_t pcg_max_iter;           /* 1000 */
  int32_t contact_sweeps;         /* 10 */
  int32_t ccd_max_iter;           /* 4096 */
  float ccd_tolerance;            /* 1e-6 */
  float ccd_min_separation_scale; /* 0.2 */
  float dcd_activation_scale;     /* 1.0 */
  int32_t mas_precision;          /* SILK_HIP_MAS_FP32 / _FP16 / _BF16: storage
                                     of the MAS bank inverses (products are
                                     accumulated in FP32 either way) */
  int32_t true_residual_stats;    /* 1 = evaluate, on the device after every
                                     PCG solve, the FP64 residual of the
                                     iterate x0 + dx and of the committed
                                     FP32 x (reported in stats) */
  int32_t pcg_criterion;          /* SILK_HIP_PCG_RR0 (reference),
                                     SILK_HIP_PCG_RHS or SILK_HIP_PCG_TRUE */
  int32_t warm_start;             /* SILK_HIP_WARM_PREVIOUS (reference),
                                     SILK_HIP_WARM_LEARNED or
                                     SILK_HIP_WARM_EXTRAPOLATE */
  int32_t warm_projection;        /* SILK_HIP_PROJECT_RESIDUAL (default) or
                                     SILK_HIP_PROJECT_GALERKIN */
  int32_t ccd_zero_toi;           /* SILK_HIP_CCD_ZERO_SKIP (reference) or
                                     SILK_HIP_CCD_ZERO_CONTACT */
  int32_t admm_graph;             /* SILK_HIP_GRAPH_OFF (launches enqueued one
                                     by one) or SILK_HIP_GRAPH_ON: each outer
                                     iteration's ADMM loop is captured into a
                                     hipGraph and launched as one graph (the
                                     executable graph is updated in place
                                     while its topology is unchanged); not
                                     used while per-kernel timing is on or a
                                     warm-start network or sample recorder is
                                     active */
  int32_t reserved;
} silk_hip_global;

enum {
  SILK_HIP_MAS_FP32 = 0,
  SILK_HIP_MAS_FP16 = 1,
  SILK_HIP_MAS_BF16 = 2
};

enum {
  /* Stop when rr <= rel^2 * rr0 (mas_cg_solver.cu:87-90). */
  SILK_HIP_PCG_RR0 = 0,
  /* Stop when ||b - A x|| <= rel * ||b|| (b the whole right-hand side of
   * the x-update, FP64 norms): a better warm start takes fewer iterations. */
  SILK_HIP_PCG_RHS = 1,
  /* As SILK_HIP_PCG_RHS, then the FP64 residual ||b - A (x0 + dx)|| of the
   * iterate (x0 + dx summed in FP64, never rounded to FP32) is tested on the
   * device; a system above rel * ||b|| restarts from its committed x, at most
   * 3 times (counted in pcg_restarts / pcg_unverified_solves). */
  SILK_HIP_PCG_TRUE = 2
};

enum {
  /* CG starts from the previous ADMM iterate (admm_solver.cu:272-292). */
  SILK_HIP_WARM_PREVIOUS = 0,
  /* CG starts from x0 + c1 * (the correction predicted by the per-vertex
   * network loaded with silk_hip_set_warm_start_net) + c2 * (the previous
   * solve's correction), (c1, c2) per cloth and coordinate from a 2x2
   * projection of the x-update system; the start is kept only when its FP64
   * residual is below that of x0. */
  SILK_HIP_WARM_LEARNED = 1,
  /* As SILK_HIP_WARM_LEARNED without the network: x0 + c2 * (the previous
   * solve's correction). */
  SILK_HIP_WARM_EXTRAPOLATE = 2
};

enum {
  /* A CCD pair already within the minimum separation ms at the segment start
   * is searched again at ms / 2; a root at t = 0 there too makes the pair no
   * collision for the segment (ccd.cpp:150-180, narrowphase.cu:68), so it
   * may cross the other primitive during the segment. */
  SILK_HIP_CCD_ZERO_SKIP = 0,
  /* As SILK_HIP_CCD_ZERO_SKIP, then a third search at zero separation: such
   * a pair collides when it would touch during the segment, so that no pair
   * separated at a segment start crosses during it. */
  SILK_HIP_CCD_ZERO_CONTACT = 1
};

enum {
  SILK_HIP_GRAPH_OFF = 0,
  SILK_HIP_GRAPH_ON = 1
};

enum {
  /* (c1, c2) minimise the FP64 residual norm ||b - A x||. */
  SILK_HIP_PROJECT_RESIDUAL = 0,
  /* (c1, c2) minimise the A-norm of the error (Galerkin). */
  SILK_HIP_PROJECT_GALERKIN = 1
};

/* Per-step counters returned by silk_hip_step. */
typedef struct {
  int32_t outer_iterations;  /* CCD segments taken */
  int32_t admm_iterations;   /* summed over segments */
  int32_t pcg_iterations;    /* summed over all PCG solves */
  int32_t pcg_solves;
  int32_t contacts;          /* active Daviet contacts at step end */
  int32_t contact_colors;
  int32_t ccd_candidates;    /* summed over segments */
  int32_t dcd_candidates;
  float min_toi;             /* earliest accepted TOI fraction (1 = none) */
  float last_pcg_rel_residual; /* sqrt(rr/rr0) of the last PCG solve */
  float max_true_rel_residual; /* max over solves of |b - A (x0 + dx)| / |b|
                                  (FP64, the PCG iterate before it is stored),
                                  when true_residual_stats is set */
  float max_committed_rel_residual; /* the same for the stored FP32 x */
  int32_t pcg_max_solve_iterations; /* largest iteration count of one solve */
  int32_t pcg_stalled_solves;       /* solves that ended at pcg_max_iter */
  int32_t warm_start_kept;   /* solves that started from the network's
                                prediction (SILK_HIP_WARM_LEARNED) */
  int32_t warm_start_solves; /* solves the network was evaluated for */
  int32_t pcg_restarts;      /* SILK_HIP_PCG_TRUE: restarts of a system's solve */
  int32_t pcg_unverified_solves; /* SILK_HIP_PCG_TRUE: solves still above the
                                    tolerance after the last restart */
  int32_t ccd_zero_toi_pairs; /* CCD candidates (summed over segments) whose
                                 last search still put a root at t = 0: no
                                 collision for their segment */
  int32_t graph_launches;     /* ADMM loops launched as a hipGraph */
  int32_t graph_instantiations; /* of which needed a new executable graph
                                   (first capture or changed topology) */
  int32_t contact_groups;     /* workgroups of the contact update when the
                                 contacts split into independent groups
                                 (cloths no contact couples), one group
                                 each; 0 when one workgroup updates all */
} silk_hip_step_stats;

/* Number of visible HIP devices (0 when none).  Never fails. */
int silk_hip_device_count(int32_t* count);

/* Create a context bound to `device`.  No device call is made until the first
 * silk_hip_step; a context without a usable device reports
 * SILK_HIP_ERR_NO_DEVICE from silk_hip_step.
 *
 * device == SILK_HIP_DEVICE_CPU creates a context of the CPU backend
 * (silk::Backend::CPU, cpu_backend.cpp:40-45 / pipeline.cpp:45-168): the same
 * scene calls, stepped by the reference's CPU algorithm (projective dynamics
 * with a prefactored global system, barrier reflection targets, CCD
 * rollback) on the host's cores (SILK_CPU_THREADS in the environment sets
 * their number).  Device-only entry points (timing, autotuning, debug
 * kernels, warm-start network) return an error on such a context.  In its
 * step statistics admm_iterations counts the PD iterations. */
int silk_hip_create(int32_t device, silk_hip_ctx** out);
int silk_hip_destroy(silk_hip_ctx* ctx);

/* Human readable description of the last error on this context ("" if none).
 * The string stays valid until the next call on the context. */
const char* silk_hip_last_error(const silk_hip_ctx* ctx);

/* World::set_global_config (cuda_backend.cu:68-85). */
int silk_hip_set_global(silk_hip_ctx* ctx, const silk_hip_global* global);

/* World::add_cloth (cuda_backend.cu:110-151).  verts: 3*nv floats, faces:
 * 3*nf ints, pins: npin vertex indices.  On an invalid mesh *handle = 0 and
 * SILK_HIP_ERR_INVALID_MESH is returned. */
int silk_hip_add_cloth(silk_hip_ctx* ctx, const silk_hip_cloth_params* cloth,
                       const silk_hip_collision_params* collision,
                       const float* verts, int64_t nv, const int32_t* faces,
                       int64_t nf, const int32_t* pins, int64_t npin,
                       uint32_t* handle);

/* World::add_obstacle (cuda_backend.cu:269-299). */
int silk_hip_add_obstacle(silk_hip_ctx* ctx,
                          const silk_hip_collision_params* collision,
                          const float* verts, int64_t nv,
                          const int32_t* faces, int64_t nf, uint32_t* handle);

/* World::remove_cloth / remove_obstacle (cuda_backend.cu:153-160,301-307).
 * is_cloth selects which kind `handle` must be. */
int silk_hip_remove(silk_hip_ctx* ctx, uint32_t handle, int32_t is_cloth);

/* World::set_cloth_config (cuda_backend.cu:197-204). */
int silk_hip_set_cloth_params(silk_hip_ctx* ctx, uint32_t handle,
                              const silk_hip_cloth_params* cloth);

/* World::set_cloth_collision_config / set_obstacle_collision_config
 * (cuda_backend.cu:206-222,309-325). */
int silk_hip_set_collision_params(silk_hip_ctx* ctx, uint32_t handle,
                                  int32_t is_cloth,
                                  const silk_hip_collision_params* collision);

/* World::set_cloth_pin_index (cuda_backend.cu:224-241). */
int silk_hip_set_pin_index(silk_hip_ctx* ctx, uint32_t handle,
                           const int32_t* pins, int64_t npin);

/* World::set_cloth_pin_position (cuda_backend.cu:243-267): 3*npin floats. */
int silk_hip_set_pin_positions(silk_hip_ctx* ctx, uint32_t handle,
                               const float* positions, int64_t n);

/* World::set_obstacle_position (cuda_backend.cu:327-352): 3*nv floats. */
int silk_hip_set_obstacle_positions(silk_hip_ctx* ctx, uint32_t handle,
                                    const float* positions, int64_t n);

/* World::solver_step (cuda_backend.cu:87-96).  stats may be NULL. */
int silk_hip_step(silk_hip_ctx* ctx, silk_hip_step_stats* stats);

/* World::solver_reset (cuda_backend.cu:98-108). */
int silk_hip_reset(silk_hip_ctx* ctx);

/* World::get_cloth_position (cuda_backend.cu:162-195): n must equal 3*nv.
 * Positions are returned in the caller's original vertex order. */
int silk_hip_get_positions(const silk_hip_ctx* ctx, uint32_t handle,
                           float* out, int64_t n);

/* ---- Introspection (no reference counterpart; used by tests and bench) ---- */

/* Vertex partition used by the MAS preconditioner for one cloth: perm maps
 * original vertex -> solver order (nv ints); offsets holds nparts+1 ints.
 * Pass NULL buffers to query sizes. */
int silk_hip_get_partition(const silk_hip_ctx* ctx, uint32_t handle,
                           int32_t* perm, int64_t nv, int32_t* offsets,
                           int64_t offsets_cap, int32_t* nparts);

/* Elapsed device time (ms) and launch count of a kernel class accumulated
 * since the last call with reset != 0.  Class 0 (whole step) is always
 * recorded; the others only while silk_hip_set_timing selects them, from
 * HIP events around each launch on the library stream: 1 = PCG solve kernel,
 * 2 = local projection, 3 = contact solve, 4 = broadphase, 5 = CCD,
 * 6 = FP64 residual; 7 = host wall time of contact assembly and colouring
 * (per call: sort, assembly, colouring, upload enqueue; no device time);
 * 8 = learned warm start (features and network). */
int silk_hip_get_timing(silk_hip_ctx* ctx, int32_t cls, double* ms,
                        int64_t* launches, int32_t reset);

/* Per-kernel event timing: 0 = off, 1 = every class, otherwise a mask of
 * the classes to time (bit c = class c; e.g. 2 = the PCG solve kernel only).
 * Events between the ADMM loop's launches keep that loop out of a hipGraph
 * (admm_graph), so a mask without classes 1, 2, 3, 6 and 8 leaves it on. */
int silk_hip_set_timing(silk_hip_ctx* ctx, int32_t on);

/* Sizes of the device scene (builds it if needed), in this order:
 * 0 cloth vertices, 1 cloth triangles, 2 off-diagonal scalar entries of the
 * PD operator, 3 MAS banks (all levels), 4 padded slots, 5 persistent PCG
 * workgroups over all systems (0 when the multi-launch PCG is used), 6
 * coarse MAS banks, 7 collider vertices, 8 persistent PCG launches per solve
 * (batches of systems whose workgroups fit the device together), 9 PCG
 * systems (cloths), 10 LDS bytes per workgroup of the persistent PCG (also
 * when it does not fit the device), 11 its variant (0: one workgroup per
 * CU, 1: two).  n: capacity of info (entries beyond it are not written). */
int silk_hip_get_scene_info(silk_hip_ctx* ctx, int64_t* info, int32_t n);

/* Collider vertex count of the current scene: cloth vertices (solver order,
 * cloths in insertion order) followed by obstacle vertices (insertion order,
 * caller order).  0 when no object pair can collide.  Builds the device
 * scene if needed. */
int silk_hip_collider_vertex_count(silk_hip_ctx* ctx, int64_t* nvt);

/* Broadphase of the current scene on explicit collider positions pos0 -> pos1
 * (n = 3 * collider vertex count floats each): triangle-point candidates as
 * (point, tri0, tri1, tri2) and edge-edge candidates as (a0, a1, b0, b1) with
 * the edge of smaller (min, max) vertex pair first, global collider ids, in
 * no particular order.  Lists are written up to their capacity (in tuples);
 * *npt / *nee receive the full counts. */
int silk_hip_debug_candidates(silk_hip_ctx* ctx, const float* pos0,
                              const float* pos1, int64_t n, int32_t* pt,
                              int64_t pt_cap, int64_t* npt, int32_t* ee,
                              int64_t ee_cap, int64_t* nee);

/* DCD on explicit positions (n = 3 * total collider vertices per end,
 * collider numbering as silk_hip_debug_candidates) at `time` with activation
 * distance dhat: 6 ints per record (type 0 PT / 1 EE, the four global vertex
 * ids in the record's order, 0) and its distance in dist.  Up to cap records
 * are copied; *nrec receives the count (dcd.cu:102-261). */
int silk_hip_debug_dcd(silk_hip_ctx* ctx, const float* pos0, const float* pos1,
                       int64_t n, float time, float dhat, int32_t* recs,
                       float* dist, int64_t cap, int64_t* nrec);

/* Discrete decisions of the last silk_hip_step: the TOI returned by the CCD
 * of every outer iteration (frame-time units, one per iteration) and the
 * number of ADMM iterations that outer iteration ran (the convergence test
 * stopped it or max_inner_iteration did), and every contact the step
 * appended, 6 ints each: the DCD call that found it (0 at the frame start, k
 * after the k-th rollback), the type and the four global vertex ids in the
 * record's order; contact_info[i] = colour << 4 | mask of that contact's
 * vertices in the contact solve (bit k: vertex k is movable and has a
 * nonzero coefficient).  NULL buffers query the counts. */
int silk_hip_debug_decisions(silk_hip_ctx* ctx, float* tois, int32_t* admm_iters,
                             int32_t cap_toi, int32_t* n_toi, int32_t* contacts,
                             int32_t* contact_info, int64_t cap_contacts,
                             int64_t* n_contacts);

/* The step's FP64 CCD kernel on n explicit queries: is_vf[i] selects
 * vertex-face (point, tri0, tri1, tri2) or edge-edge (a0, a1, b0, b1);
 * p0 / p1 hold 12 doubles per query, err 3 per query.  hit[i] = 1 and toi[i]
 * (local time in [0, 1]) for a hit, hit[i] = 0 otherwise.  A root at t = 0
 * is refined as silk_hip_global.ccd_zero_toi selects. */
int silk_hip_debug_ccd(silk_hip_ctx* ctx, int64_t n, const int32_t* is_vf,
                       const double* p0, const double* p1, const double* err,
                       double ms, double tolerance, int32_t max_iter,
                       int32_t* hit, double* toi);

/* The step's earliest-TOI CCD kernel on the same explicit queries: the
 * smallest toi over the hits with toi > 0, or 1 when there is none (the
 * reduction of find_collision.cu:142-190 and narrowphase.cu:32-80 over one
 * query list; queries that cannot lower the running minimum are cut short,
 * which leaves the minimum unchanged). */
int silk_hip_debug_ccd_min(silk_hip_ctx* ctx, int64_t n, const int32_t* is_vf,
                           const double* p0, const double* p1,
                           const double* err, double ms, double tolerance,
                           int32_t max_iter, double* min_toi);

/* Capacity of silk_hip_ccd_min_noalloc (queries per call). */
#define SILK_HIP_CCD_NOALLOC_MAX 16384

/* No counterpart in the reference: silk_hip_debug_ccd_min for callers that
 * cannot allocate device memory (code under HIP graph capture, tight query
 * loops, harnesses that forbid hipMalloc).  Same arguments and result; the
 * queries are staged and the running minimum kept in a fixed device array
 * (about 3.5 MB per device, held until the process ends) instead of buffers
 * allocated per call.  n above SILK_HIP_CCD_NOALLOC_MAX is rejected with
 * SILK_HIP_ERR_INVALID_CONFIG.  Every call on a device shares that device's
 * array: calls that overlap on one device (from several threads or
 * contexts) race. */
int silk_hip_ccd_min_noalloc(silk_hip_ctx* ctx, int64_t n, const int32_t* is_vf,
                             const double* p0, const double* p1, const double* err,
                             double ms, double tolerance, int32_t max_iter,
                             double* min_toi);

/* Device copy of `bytes` (a multiple of 4) from one buffer to another, reps
 * times, with the library's copy kernel (k_copy_unless); *ms receives the
 * elapsed device time.  A reference of known traffic for hardware counters:
 * each launch reads and writes `bytes`. */
int silk_hip_debug_copy(silk_hip_ctx* ctx, int64_t bytes, int32_t reps, double* ms);

/* One PCG solve of the scene's x-update operator K + diag(d) on an explicit
 * right-hand side, from x = 0, with the current global configuration
 * (criterion, tolerances, mas_precision): d (nv floats) and b (3 nv) in solver
 * order (cloths concatenated, each in the order of silk_hip_get_partition);
 * rel_tol replaces the adaptive tolerance.  x (3 nv) receives the solution,
 * *iterations the iterations summed over the cloths' systems. */
int silk_hip_debug_pcg(silk_hip_ctx* ctx, const float* d, int64_t nv,
                       const float* b, int64_t n3, float rel_tol, float* x,
                       int32_t* iterations);

/* The contact solve's Coulomb projection (replaces solve_coulomb_contact,
 * contact_constraints.cu:391-411, as the contact kernels run it) on n
 * explicit cases: u and normal hold 3 floats per case, mu 1, out 3. */
int silk_hip_debug_coulomb(silk_hip_ctx* ctx, int64_t n, const float* u,
                           const float* normal, const float* mu, float* out);

/* ---- Learned CG warm start ----
 * No counterpart in the reference (it starts every CG solve from the previous
 * ADMM iterate, mas_cg_solver.cu:193-311).  For callers that solve with
 * SILK_HIP_PCG_RHS, where a better start takes fewer iterations.  A per-vertex
 * network (16 features of the vertex and its operator 1-ring -> 64 -> 64 -> 1,
 * SiLU, BF16 MFMA with FP32 accumulation) predicts each x-update's
 * correction; the prediction is used only when its FP64 residual is below the
 * previous iterate's, and PCG then runs to the same stopping test. */

/* Parameters of the network (FP32, in this order): W1[64][16], b1[64],
 * W2[64][64], b2[64], W3[64], b3. */
#define SILK_HIP_WARM_NET_PARAMS 5313

/* Load the network; n must equal SILK_HIP_WARM_NET_PARAMS. */
int silk_hip_set_warm_start_net(silk_hip_ctx* ctx, const float* params,
                                int64_t n);

/* Read the network back; SILK_HIP_ERR_NEED_INIT_SOLVER when none is loaded. */
int silk_hip_get_warm_start_net(silk_hip_ctx* ctx, float* params, int64_t n);

/* Record training samples during the following silk_hip_step calls: each
 * (vertex, coordinate) of each x-update solve (started from the previous
 * iterate) is kept with probability `rate`, up to `capacity` samples held in
 * device memory; earlier samples are dropped.  capacity 0 stops recording
 * and frees them. */
int silk_hip_warm_start_record(silk_hip_ctx* ctx, int64_t capacity, float rate);

/* Number of samples recorded so far. */
int silk_hip_warm_start_samples(silk_hip_ctx* ctx, int64_t* n);

/* Train on the recorded samples on the device: `steps` Adam steps with
 * learning rate lr on batches of `batch` samples drawn by a hash of (step,
 * index); starts from the loaded network, or from a seeded initialisation
 * when none is loaded.  loss (optional, `steps` floats) receives the batch
 * losses (mean squared error of the correction / sigma). */
int silk_hip_warm_start_train(silk_hip_ctx* ctx, int32_t steps, int32_t batch,
                              float lr, uint64_t seed, float* loss);

/* Forward and backward of the loaded network on n explicit samples (16
 * features each, rounded to BF16 on the device) with targets: y (n), the
 * gradient of the mean squared error (SILK_HIP_WARM_NET_PARAMS) and the
 * loss.  For gradient checks. */
int silk_hip_debug_warm_net(silk_hip_ctx* ctx, int64_t n, const float* feat,
                            const float* target, float* y, float* grad,
                            double* loss);

/* ---- Launch autotuning ----
 * No counterpart in the reference (its launch shapes are fixed at compile
 * time).  For callers that step one scene many times: times, on the current
 * scene, the broadphase at traversal split depths 2..7 (5 launches each) and
 * the persistent PCG with one and with two workgroups per CU (a solve of the
 * scene's matrix from x = 0 to 1e-6 on a fixed right-hand side, best of 3),
 * and keeps the fastest of each until the scene is rebuilt.  Every choice
 * gives the same results; only the time changes.  Needs one silk_hip_step on
 * the current scene first (SILK_HIP_ERR_INVALID_CONFIG otherwise).
 * report (cap >= SILK_HIP_AUTOTUNE_REPORT int32) receives: [0] the chosen
 * split depth, [1] the chosen PCG variant (0 one workgroup per CU, 1 two,
 * -1 no persistent PCG), [2..7] microseconds per broadphase launch at depth
 * 2..7, [8..9] microseconds per PCG solve of variants 0 and 1 (-1 where the
 * variant does not fit the device). */
#define SILK_HIP_AUTOTUNE_REPORT 10
int silk_hip_autotune(silk_hip_ctx* ctx, int32_t* report, int32_t cap);

#ifdef __cplusplus
}
#endif

#endif /* SILK_HIP_H_ */
