// Public C++ API of the cloth solver.
//
// Declaration-identical to the reference interface
// (silk/include/silk/silk.hpp:16-239): the same structs, defaults, member
// functions and argument meaning, so an application written against silk
// links against this library without source changes.  Backend::GPU selects the
// MI355X (HIP) projective-dynamics step; Backend::CPU is not provided by this
// library and init(Backend::CPU) reports InvalidConfig.
#pragma once

#include <cstdint>
#include <memory>
#include <span>
#include <vector>

#include "silk/result.hpp"

namespace silk {

/// Triangle mesh.  verts = [x0,y0,z0, x1,y1,z1, ...], faces = [i0,j0,k0, ...].
/// The spans are read during the call only; their contents are copied.
struct MeshConfig {
  std::span<const float> verts;
  std::span<const int> faces;
};

/// Collision settings of one object.
struct CollisionConfig {
  bool is_collision_on = true;
  bool is_self_collision_on = true;
  int group = 0;
  float restitution = 0.3f;
  float friction = 0.3f;
};

/// Material of one cloth.
struct ClothConfig {
  /// Membrane (in-plane) stiffness; also the ADMM penalty.
  float elastic_stiffness = 100.0f;
  /// Quadratic bending stiffness.
  float bending_stiffness = 0.0001f;
  /// Mass per unit area.
  float density = 0.1f;
  /// Fraction of velocity removed at the start of every step.
  float damping = 0.01f;
};

/// Scene-wide solver settings.
struct GlobalConfig {
  float acceleration_x = 0.0f;
  float acceleration_y = 0.0f;
  float acceleration_z = -10.0f;

  /// Cap on CCD rollback segments per step.
  int max_outer_iteration = 100;
  /// Cap on ADMM iterations per segment.
  int max_inner_iteration = 100;

  float dt = 1.0f / 60.0f;

  float linear_solver_abs_tol = 1e-7f;
  float linear_solver_rel_tol_min = 1e-6f;
  float linear_solver_rel_tol_max = 1e-3f;
  float initial_linear_rel_tol = 1e-2f;

  float admm_abs_tol = 1e-5f;
  float admm_rel_tol = 1e-3f;
};

enum class Backend { CPU, GPU };

/// Simulation world.  Move-only; not thread safe.
class World {
 private:
  class Impl;
  std::unique_ptr<Impl> impl_;

 public:
  World();
  ~World();
  World(const World&) = delete;
  World(World&&);
  World& operator=(const World&) = delete;
  World& operator=(World&&);

  // Global API
  [[nodiscard]] Result init(Backend backend);
  [[nodiscard]] Result set_global_config(GlobalConfig config);
  void clear();

  // Solver API
  [[nodiscard]] Result solver_step();
  [[nodiscard]] Result solver_reset();

  // Cloth API
  [[nodiscard]] Result add_cloth(ClothConfig cloth_config,
                                 CollisionConfig collision_config,
                                 MeshConfig mesh_config,
                                 std::span<const int> pin_index,
                                 uint32_t& handle);
  [[nodiscard]] Result remove_cloth(uint32_t handle);
  [[nodiscard]] Result get_cloth_position(uint32_t handle,
                                          std::span<float> position) const;
  [[nodiscard]] Result set_cloth_config(uint32_t handle, ClothConfig config);
  [[nodiscard]] Result set_cloth_collision_config(uint32_t handle,
                                                  CollisionConfig config);
  [[nodiscard]] Result set_cloth_pin_index(uint32_t handle,
                                           std::span<const int> pin_index);
  [[nodiscard]] Result set_cloth_pin_position(uint32_t handle,
                                              std::span<const float> position);

  // Obstacle API
  [[nodiscard]] Result add_obstacle(CollisionConfig collision_config,
                                    MeshConfig mesh_config, uint32_t& handle);
  [[nodiscard]] Result remove_obstacle(uint32_t handle);
  [[nodiscard]] Result set_obstacle_collision_config(uint32_t handle,
                                                     CollisionConfig config);
  [[nodiscard]] Result set_obstacle_position(uint32_t handle,
                                             std::span<const float> position);
};

}  // namespace silk
