// World API driver used by tests/test_world_api.py.
//
// "host" mode needs no GPU: it checks the front-end contract (NotInitialized
// before init, InvalidConfig for the CPU backend, NoCudaSupport without a
// device, Result::to_string formatting).  "gpu" mode adds a cloth through
// World, steps it and checks the reference's validation codes
// (cuda_backend.cu:110-352).  Prints one "PASS <name>" / "FAIL <name>" line
// per check and exits non-zero on any failure.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "silk/silk.hpp"

namespace {

int failures = 0;

void check(bool ok, const char* name) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", name);
  if (!ok) ++failures;
}

void grid(int n, std::vector<float>& V, std::vector<int>& F) {
  V.clear();
  F.clear();
  for (int j = 0; j <= n; ++j)
    for (int i = 0; i <= n; ++i) {
      V.push_back(static_cast<float>(i) / n);
      V.push_back(static_cast<float>(j) / n);
      V.push_back(0.0f);
    }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      int a = j * (n + 1) + i, b = a + 1, c = a + n + 1, d = c + 1;
      F.insert(F.end(), {a, b, c, c, b, d});
    }
}

}  // namespace

int main(int argc, char** argv) {
  const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
  silk::World w;
  silk::Result r = w.solver_step();
  check(!r && r.code == silk::ErrorCode::NotInitialized, "step_before_init");
  check(r.to_string() == "NotInitialized", "to_string_no_detail");
  check(silk::Result::ok().to_string() == "OK", "to_string_ok");
  auto e = silk::Result::error(silk::ErrorCode::InvalidPin, "bad");
  check(e.to_string() == "InvalidPin. Reason: bad.", "to_string_detail");
  r = w.init(silk::Backend::CPU);
  check(!r && r.code == silk::ErrorCode::InvalidConfig, "cpu_backend_rejected");

  r = w.init(silk::Backend::GPU);
  if (!gpu) {
    // Without a device the GPU backend must report NoCudaSupport; with one it
    // must succeed.  Host mode accepts either but checks consistency.
    check(r || r.code == silk::ErrorCode::NoCudaSupport, "gpu_init_consistent");
    return failures == 0 ? 0 : 1;
  }
  check(static_cast<bool>(r), "gpu_init");
  silk::GlobalConfig g;
  g.acceleration_z = -9.81f;
  g.max_inner_iteration = 10;
  check(static_cast<bool>(w.set_global_config(g)), "set_global_config");

  std::vector<float> V;
  std::vector<int> F;
  grid(8, V, F);
  silk::CollisionConfig cc;
  cc.is_collision_on = false;
  uint32_t h = 123;
  std::vector<int> pins = {8 * 9, 8 * 9 + 8};

  // Invalid mesh: NaN vertex.
  std::vector<float> Vbad = V;
  Vbad[4] = NAN;
  r = w.add_cloth({}, cc, {Vbad, F}, pins, h);
  check(!r && r.code == silk::ErrorCode::InvalidMesh && h == 0, "nan_mesh");
  // Invalid pin.
  std::vector<int> badpins = {1000};
  r = w.add_cloth({}, cc, {V, F}, badpins, h);
  check(!r && r.code == silk::ErrorCode::InvalidPin, "bad_pin");

  r = w.add_cloth({}, cc, {V, F}, pins, h);
  check(r && h != 0, "add_cloth");
  std::vector<float> pos(V.size());
  r = w.get_cloth_position(h, pos);
  check(r && pos == V, "initial_position");
  std::vector<float> small(3);
  check(w.get_cloth_position(h, small).code == silk::ErrorCode::InvalidPosition,
        "position_size");
  check(w.get_cloth_position(h + 1, pos).code == silk::ErrorCode::InvalidHandle,
        "bad_handle");
  for (int s = 0; s < 5; ++s) {
    r = w.solver_step();
    if (!r) break;
  }
  if (!r) std::printf("step error: %s\n", r.to_string().c_str());
  check(static_cast<bool>(r), "steps");
  r = w.get_cloth_position(h, pos);
  bool fell = r && pos[2 * 3 + 2] < -1e-4f;  // unpinned bottom vertex falls
  bool pinned = r && std::fabs(pos[3 * pins[0] + 2]) < 1e-6f;
  check(fell, "cloth_falls");
  check(pinned, "pin_holds");
  check(static_cast<bool>(w.solver_reset()), "reset");
  r = w.get_cloth_position(h, pos);
  check(r && pos == V, "reset_restores_initial");
  check(static_cast<bool>(w.remove_cloth(h)), "remove_cloth");
  check(w.remove_cloth(h).code == silk::ErrorCode::InvalidHandle,
        "remove_twice");
  return failures == 0 ? 0 : 1;
}
