// Collision subsystem of the step: colliders, broadphase, DCD contacts, CCD
// and the Daviet contact projection (reference: collision/*.cu,
// solver/contact_constraints.cu, solver/contact_coloring.cu).
#pragma once

#include <memory>
#include <vector>

#include "silk_hip.h"

namespace silk::amd {

class Context;
struct Scene;

/// Device pointers the ADMM kernels need from the contact set.  All null when
/// no contact is active.
struct ContactView {
  const float* w = nullptr;     // nv: active per-vertex contact weight
  const float* p = nullptr;     // 3*nv: projected position
  const float* eta = nullptr;   // 3*nv: multiplier
  const float* dual = nullptr;  // 3*nv: w (p - p_prev)
};

class CollisionWorld {
 public:
  virtual ~CollisionWorld() = default;
  /// Frame start: refit obstacles and cloth colliders, minimum separation and
  /// activation distance, initial DCD contact set (main_loop.cu:186-216).
  virtual void begin_step(const float* outer_x, float* ms, float* dhat,
                          silk_hip_step_stats* stats) = 0;
  /// Earliest TOI of the segment outer_x -> inner_x in frame-time units,
  /// `interval` when none (find_collision.cu:140-190).
  virtual float min_toi(const float* outer_x, const float* inner_x,
                        float time_start, float interval, float ms,
                        silk_hip_step_stats* stats) = 0;
  /// After a rollback: refit at outer_x and append DCD contacts at `time`
  /// (main_loop.cu:267-277).
  virtual void rollback_contacts(const float* outer_x, float time, float dhat,
                                 silk_hip_step_stats* stats) = 0;
  /// ADMM contact update for the current x (contact_constraints.cu:246-264).
  /// skip: optional device flag; the update is a no-op while *skip != 0.
  virtual void contact_update(const float* x, const int* skip) = 0;
  virtual ContactView contact_view() const = 0;
  virtual int active_dof() const = 0;
  virtual void end_step(silk_hip_step_stats* stats) = 0;
  /// Broadphase on explicit host positions (3 * total collider vertices per
  /// end): PT tuples (point, tri0, tri1, tri2) and EE tuples (a0, a1, b0, b1)
  /// in canonical edge order, global vertex ids.
  virtual void debug_candidates(const float* pos0, const float* pos1,
                                std::vector<int>& pt, std::vector<int>& ee) = 0;
  virtual int total_vertices() const = 0;
  /// DCD on explicit host positions (3 * total collider vertices per end) at
  /// `time` with activation distance dhat: 6 ints per record (type, then the
  /// four global vertex ids in the record's order) and its distance.
  virtual void debug_dcd(const float* pos0, const float* pos1, float time,
                         float dhat, std::vector<int>& recs,
                         std::vector<float>& dist) = 0;
  /// Contacts appended during the last step, 6 ints each: the DCD call that
  /// found it (0 at the frame start, k after the k-th rollback), the type and
  /// the four global vertex ids in the record's order.
  virtual const std::vector<int>& decision_log() const = 0;
  /// Per appended contact: its colour << 4 | the mask of its vertices that
  /// take part in the contact solve (nonzero coefficient on a movable
  /// vertex; bit k for vertex k).
  virtual const std::vector<int>& decision_info() const = 0;
  /// Broadphase work split: each query is walked from every subtree at this
  /// depth of the trees (capped by a tree's height).  Candidate sets do not
  /// depend on it, only their order and the launch's speed.
  virtual int split_depth() const = 0;
  virtual void set_split_depth(int depth) = 0;
  /// Mean device time (ms) of `reps` broadphase launches on the current
  /// collider positions.
  virtual float time_broadphase(int reps) = 0;
};

/// nullptr when no object has collision enabled.
std::unique_ptr<CollisionWorld> make_collision_world(Context& ctx, Scene& S);

}  // namespace silk::amd
