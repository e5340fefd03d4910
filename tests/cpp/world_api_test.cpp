// World API driver used by tests/test_world_api.py.
//
// "host" mode needs no GPU: it checks the front-end contract (NotInitialized
// before init, NoCudaSupport without a device, Result::to_string formatting)
// and runs the scene checks below on Backend::CPU.  "gpu" mode runs them on
// Backend::GPU too: add a cloth through World, step it and check the
// reference's validation codes (cuda_backend.cu:110-352).  Prints one
// "PASS <name>" / "FAIL <name>" line per check and exits non-zero on any
// failure.
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

// Cloth validation, stepping, pins, reset and removal through World on an
// initialised backend.  Pins of the CPU backend are penalty terms
// (cloth_solver_context.cpp:17), so a pinned vertex may move by about its
// share of the weight over the penalty; the GPU's are exact.
void scene_checks(silk::World& w, const std::string& tag, float pin_tol) {
  auto name = [&](const char* n) { return (tag + "_" + n); };
  silk::GlobalConfig g;
  g.acceleration_z = -9.81f;
  g.max_inner_iteration = 10;
  check(static_cast<bool>(w.set_global_config(g)), name("set_global_config").c_str());

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
  silk::Result r = w.add_cloth({}, cc, {Vbad, F}, pins, h);
  check(!r && r.code == silk::ErrorCode::InvalidMesh && h == 0, name("nan_mesh").c_str());
  // Invalid pin.
  std::vector<int> badpins = {1000};
  r = w.add_cloth({}, cc, {V, F}, badpins, h);
  check(!r && r.code == silk::ErrorCode::InvalidPin, name("bad_pin").c_str());

  r = w.add_cloth({}, cc, {V, F}, pins, h);
  check(r && h != 0, name("add_cloth").c_str());
  std::vector<float> pos(V.size());
  r = w.get_cloth_position(h, pos);
  check(r && pos == V, name("initial_position").c_str());
  std::vector<float> small(3);
  check(w.get_cloth_position(h, small).code == silk::ErrorCode::InvalidPosition,
        name("position_size").c_str());
  check(w.get_cloth_position(h + 1, pos).code == silk::ErrorCode::InvalidHandle,
        name("bad_handle").c_str());
  for (int s = 0; s < 5; ++s) {
    r = w.solver_step();
    if (!r) break;
  }
  if (!r) std::printf("step error: %s\n", r.to_string().c_str());
  check(static_cast<bool>(r), name("steps").c_str());
  r = w.get_cloth_position(h, pos);
  bool fell = r && pos[2 * 3 + 2] < -1e-4f;  // unpinned bottom vertex falls
  bool pinned = r && std::fabs(pos[3 * pins[0] + 2]) <= pin_tol;
  check(fell, name("cloth_falls").c_str());
  check(pinned, name("pin_holds").c_str());
  check(static_cast<bool>(w.solver_reset()), name("reset").c_str());
  r = w.get_cloth_position(h, pos);
  check(r && pos == V, name("reset_restores_initial").c_str());
  check(static_cast<bool>(w.remove_cloth(h)), name("remove_cloth").c_str());
  check(w.remove_cloth(h).code == silk::ErrorCode::InvalidHandle,
        name("remove_twice").c_str());
}

int main(int argc, char** argv) {
  const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
  silk::World w;
  silk::Result r = w.solver_step();
  check(!r && r.code == silk::ErrorCode::NotInitialized, "step_before_init");
  check(r.to_string() == "NotInitialized", "to_string_no_detail");
  check(silk::Result::ok().to_string() == "OK", "to_string_ok");
  auto e = silk::Result::error(silk::ErrorCode::InvalidPin, "bad");
  check(e.to_string() == "InvalidPin. Reason: bad.", "to_string_detail");
  {
    silk::World wc;
    r = wc.init(silk::Backend::CPU);
    check(static_cast<bool>(r), "cpu_init");
    if (r) scene_checks(wc, "cpu", 1e-3f);
  }

  r = w.init(silk::Backend::GPU);
  if (!gpu) {
    // Without a device the GPU backend must report NoCudaSupport; with one it
    // must succeed.  Host mode accepts either but checks consistency.
    check(r || r.code == silk::ErrorCode::NoCudaSupport, "gpu_init_consistent");
    return failures == 0 ? 0 : 1;
  }
  check(static_cast<bool>(r), "gpu_init");
  scene_checks(w, "gpu", 1e-6f);
  return failures == 0 ? 0 : 1;
}
