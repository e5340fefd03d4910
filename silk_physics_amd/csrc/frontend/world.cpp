// silk::World and silk::Result on top of the C ABI (include/silk_hip.h).
//
// World keeps the reference pimpl + backend-interface shape
// (silk/src/common/world.cpp:15-166, common/backend.hpp:9-45): every call
// checks that a backend exists, then forwards.  Both backends go through
// HipBackend, a thin mapping from std::span arguments to the C ABI and from
// C status codes back to silk::ErrorCode; Backend::CPU opens a context of
// the library's CPU solver (device SILK_HIP_DEVICE_CPU).
#include <cstdio>
#include <memory>
#include <span>
#include <string>

#include "silk/silk.hpp"
#include "silk_hip.h"

namespace silk {

// ------------------------------------------------------------------ Result

namespace {
const char* code_name(ErrorCode c) {
  switch (c) {
    case ErrorCode::Unknown: return "Unknown";
    case ErrorCode::NotInitialized: return "NotInitialized";
    case ErrorCode::InvalidConfig: return "InvalidConfig";
    case ErrorCode::TooManyBody: return "TooManyBody";
    case ErrorCode::InvalidHandle: return "InvalidHandle";
    case ErrorCode::InvalidMesh: return "InvalidMesh";
    case ErrorCode::InvalidPin: return "InvalidPin";
    case ErrorCode::InvalidPosition: return "InvalidPosition";
    case ErrorCode::CholeskyDecompositionFail: return "CholeskyDecompositionFail";
    case ErrorCode::NeedInitSolverFirst: return "NeedInitSolverFirst";
    case ErrorCode::IterativeSolveFail: return "IterativeSolveFail";
    case ErrorCode::NoCudaSupport: return "NoCudaSupport";
  }
  return "Unknown";
}
}  // namespace

Result Result::ok() { return Result(); }

Result Result::error(ErrorCode error_code, std::string_view detail) {
  Result r;
  r.success = false;
  r.code = error_code;
  r.detail = std::string(detail);
  return r;
}

Result::operator bool() const { return success; }

std::string Result::to_string() const {
  if (success) return "OK";
  if (detail.empty()) return code_name(code);
  return std::string(code_name(code)) + ". Reason: " + detail + ".";
}

// -------------------------------------------------------------- HipBackend

namespace {

silk_hip_cloth_params to_c(const ClothConfig& c) {
  return {c.elastic_stiffness, c.bending_stiffness, c.density, c.damping};
}
silk_hip_collision_params to_c(const CollisionConfig& c) {
  return {c.is_collision_on ? 1 : 0, c.is_self_collision_on ? 1 : 0, c.group,
          c.restitution, c.friction};
}

class HipBackend {
 public:
  explicit HipBackend(silk_hip_ctx* ctx) : ctx_(ctx) {}
  ~HipBackend() { silk_hip_destroy(ctx_); }
  HipBackend(const HipBackend&) = delete;
  HipBackend& operator=(const HipBackend&) = delete;

  Result wrap(int status) const {
    if (status == SILK_HIP_OK) return Result::ok();
    return Result::error(static_cast<ErrorCode>(status - 1),
                         silk_hip_last_error(ctx_));
  }

  Result set_global_config(const GlobalConfig& c) {
    silk_hip_global g{};
    g.acceleration[0] = c.acceleration_x;
    g.acceleration[1] = c.acceleration_y;
    g.acceleration[2] = c.acceleration_z;
    g.max_outer_iteration = c.max_outer_iteration;
    g.max_inner_iteration = c.max_inner_iteration;
    g.dt = c.dt;
    g.linear_solver_abs_tol = c.linear_solver_abs_tol;
    g.linear_solver_rel_tol_min = c.linear_solver_rel_tol_min;
    g.linear_solver_rel_tol_max = c.linear_solver_rel_tol_max;
    g.initial_linear_rel_tol = c.initial_linear_rel_tol;
    g.admm_abs_tol = c.admm_abs_tol;
    g.admm_rel_tol = c.admm_rel_tol;
    return wrap(silk_hip_set_global(ctx_, &g));
  }

  silk_hip_ctx* ctx_;
};

}  // namespace

class World::Impl {
 public:
  std::unique_ptr<HipBackend> backend;
};

World::World() : impl_(std::make_unique<Impl>()) {}
World::~World() = default;
World::World(World&&) = default;
World& World::operator=(World&&) = default;

Result World::init(Backend backend) {
  switch (backend) {
    case Backend::CPU: {
      // The reference's CPU solver on the host's cores, behind the same C
      // ABI (cpu_backend.cpp:40-45).
      silk_hip_ctx* ctx = nullptr;
      if (silk_hip_create(SILK_HIP_DEVICE_CPU, &ctx) != SILK_HIP_OK) {
        return Result::error(ErrorCode::Unknown, "context creation failed");
      }
      impl_->backend = std::make_unique<HipBackend>(ctx);
      return Result::ok();
    }
    case Backend::GPU: {
      int32_t n = 0;
      silk_hip_device_count(&n);
      if (n <= 0) {
        return Result::error(ErrorCode::NoCudaSupport,
                             "HIP runtime/device not available");
      }
      silk_hip_ctx* ctx = nullptr;
      if (silk_hip_create(0, &ctx) != SILK_HIP_OK) {
        return Result::error(ErrorCode::Unknown, "context creation failed");
      }
      impl_->backend = std::make_unique<HipBackend>(ctx);
      return Result::ok();
    }
  }
  return Result::error(ErrorCode::InvalidConfig, "Unknown backend");
}

#define SILK_REQUIRE_BACKEND()                      \
  if (!impl_->backend) {                            \
    return Result::error(ErrorCode::NotInitialized); \
  }                                                 \
  HipBackend& be = *impl_->backend

Result World::set_global_config(GlobalConfig config) {
  SILK_REQUIRE_BACKEND();
  return be.set_global_config(config);
}

void World::clear() { impl_->backend.reset(); }

Result World::solver_step() {
  SILK_REQUIRE_BACKEND();
  return be.wrap(silk_hip_step(be.ctx_, nullptr));
}

Result World::solver_reset() {
  SILK_REQUIRE_BACKEND();
  return be.wrap(silk_hip_reset(be.ctx_));
}

Result World::add_cloth(ClothConfig cloth_config,
                        CollisionConfig collision_config,
                        MeshConfig mesh_config, std::span<const int> pin_index,
                        uint32_t& handle) {
  SILK_REQUIRE_BACKEND();
  const silk_hip_cloth_params cp = to_c(cloth_config);
  const silk_hip_collision_params cc = to_c(collision_config);
  if (mesh_config.verts.size() % 3 != 0 || mesh_config.faces.size() % 3 != 0) {
    handle = 0;
    return Result::error(ErrorCode::InvalidMesh);
  }
  return be.wrap(silk_hip_add_cloth(
      be.ctx_, &cp, &cc, mesh_config.verts.data(),
      static_cast<int64_t>(mesh_config.verts.size() / 3),
      mesh_config.faces.data(),
      static_cast<int64_t>(mesh_config.faces.size() / 3), pin_index.data(),
      static_cast<int64_t>(pin_index.size()), &handle));
}

Result World::remove_cloth(uint32_t handle) {
  SILK_REQUIRE_BACKEND();
  return be.wrap(silk_hip_remove(be.ctx_, handle, 1));
}

Result World::get_cloth_position(uint32_t handle,
                                 std::span<float> position) const {
  if (!impl_->backend) return Result::error(ErrorCode::NotInitialized);
  const HipBackend& be = *impl_->backend;
  return be.wrap(silk_hip_get_positions(be.ctx_, handle, position.data(),
                                        static_cast<int64_t>(position.size())));
}

Result World::set_cloth_config(uint32_t handle, ClothConfig config) {
  SILK_REQUIRE_BACKEND();
  const silk_hip_cloth_params cp = to_c(config);
  return be.wrap(silk_hip_set_cloth_params(be.ctx_, handle, &cp));
}

Result World::set_cloth_collision_config(uint32_t handle,
                                         CollisionConfig config) {
  SILK_REQUIRE_BACKEND();
  const silk_hip_collision_params cc = to_c(config);
  return be.wrap(silk_hip_set_collision_params(be.ctx_, handle, 1, &cc));
}

Result World::set_cloth_pin_index(uint32_t handle,
                                  std::span<const int> pin_index) {
  SILK_REQUIRE_BACKEND();
  return be.wrap(silk_hip_set_pin_index(be.ctx_, handle, pin_index.data(),
                                        static_cast<int64_t>(pin_index.size())));
}

Result World::set_cloth_pin_position(uint32_t handle,
                                     std::span<const float> position) {
  SILK_REQUIRE_BACKEND();
  return be.wrap(silk_hip_set_pin_positions(
      be.ctx_, handle, position.data(), static_cast<int64_t>(position.size())));
}

Result World::add_obstacle(CollisionConfig collision_config,
                           MeshConfig mesh_config, uint32_t& handle) {
  SILK_REQUIRE_BACKEND();
  const silk_hip_collision_params cc = to_c(collision_config);
  if (mesh_config.verts.size() % 3 != 0 || mesh_config.faces.size() % 3 != 0) {
    handle = 0;
    return Result::error(ErrorCode::InvalidMesh);
  }
  return be.wrap(silk_hip_add_obstacle(
      be.ctx_, &cc, mesh_config.verts.data(),
      static_cast<int64_t>(mesh_config.verts.size() / 3),
      mesh_config.faces.data(),
      static_cast<int64_t>(mesh_config.faces.size() / 3), &handle));
}

Result World::remove_obstacle(uint32_t handle) {
  SILK_REQUIRE_BACKEND();
  return be.wrap(silk_hip_remove(be.ctx_, handle, 0));
}

Result World::set_obstacle_collision_config(uint32_t handle,
                                            CollisionConfig config) {
  SILK_REQUIRE_BACKEND();
  const silk_hip_collision_params cc = to_c(config);
  return be.wrap(silk_hip_set_collision_params(be.ctx_, handle, 0, &cc));
}

Result World::set_obstacle_position(uint32_t handle,
                                    std::span<const float> position) {
  SILK_REQUIRE_BACKEND();
  return be.wrap(silk_hip_set_obstacle_positions(
      be.ctx_, handle, position.data(), static_cast<int64_t>(position.size())));
}

}  // namespace silk
