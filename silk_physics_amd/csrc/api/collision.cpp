// Collision subsystem construction.
#include "api/collision.hpp"

#include "api/context.hpp"

namespace silk::amd {

std::unique_ptr<CollisionWorld> make_collision_world(Context& ctx, Scene& S) {
  (void)ctx;
  (void)S;
  return nullptr;
}

}  // namespace silk::amd
