// The CPU backend (silk::Backend::CPU): the reference's CPU solver step on
// the host's cores.
//
// Algorithm of silk/src/backend/cpu/solver/pipeline.cpp:45-168: projective
// dynamics with a prefactored global system, an outer loop over CCD segments
// (barrier reflection targets from the DCD contacts, rollback to 0.8 of the
// earliest time of impact) and an inner loop of per-face stretch projections
// and global solves until the update falls below 5 % of the scene size.
// Host-native choices: a project-written nested-dissection Cholesky with
// partial refactoring for barrier changes (the reference: CHOLMOD with rank
// updates), a uniform-grid broadphase (the reference: SAP over objects and
// KD-trees over primitives), an FP64 interval root finder for the CCD, and a
// persistent worker pool (the reference: TBB).
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "host/geometry.hpp"
#include "silk_hip.h"

namespace silk::amd::cpu {

struct ClothInput {
  const TriMesh* mesh = nullptr;
  silk_hip_cloth_params cloth{};
  silk_hip_collision_params coll{};
  std::vector<int> pins;
  std::vector<float> pin_pos;  // 3 per pin
  std::vector<float> x, v;     // state, 3 per vertex
};

struct ObstacleInput {
  const TriMesh* mesh = nullptr;
  silk_hip_collision_params coll{};
};

struct StepStats {
  int outer = 0;            // CCD segments
  int inner = 0;            // PD iterations (summed over segments)
  int collisions = 0;       // DCD contacts (summed over segments)
  int ccd_candidates = 0;   // broadphase pairs tested by the CCD
  float min_toi = 1.0f;     // earliest accepted time of impact
  int rows_refactored = 0;  // Cholesky rows recomputed for barrier changes
};

class CpuWorld {
 public:
  /// threads <= 0: the host's hardware concurrency.
  CpuWorld(const silk_hip_global& g, int threads);
  ~CpuWorld();
  CpuWorld(const CpuWorld&) = delete;
  CpuWorld& operator=(const CpuWorld&) = delete;

  /// Throws std::runtime_error when the system matrix is not positive
  /// definite.
  void add_cloth(const ClothInput& c);
  void add_obstacle(const ObstacleInput& o);
  /// Obstacle positions at the start and the end of the coming step.
  void set_obstacle_frame(int idx, const float* prev, const float* curr);
  void set_pin_targets(int cloth, const float* pos);
  /// One step; false when a factorisation failed or the state became
  /// non-finite (the state is then left as it was).
  bool step(StepStats* st);

  const std::vector<float>& x(int cloth) const;
  const std::vector<float>& v(int cloth) const;
  int threads() const;

  struct Impl;

 private:
  std::unique_ptr<Impl> p_;
};

}  // namespace silk::amd::cpu
