// Headless driver: runs a scene file in the reference's JSON format through
// silk::World and writes the trajectory (demo/src/headless.cpp:402-502,
// json_parse.cpp, docs/config_format.md of the reference).
//
//   silk_headless --config scene.json [--output out.silktraj] [--bench]
//                 [--steps N] [--backend gpu|cpu] [--quiet] [--check]
//
// --check parses the scene and loads its meshes, prints one JSON line per
// object (vertices, triangles, pins, group, transformed bounding box) and
// exits without touching a device.
//
// Scene keys and defaults follow the reference's config.hpp: "global" (dt,
// max_outer_iteration, max_inner_iteration, acceleration, total_steps, the
// linear-solver and ADMM tolerances) and "objects" (type cloth/obstacle,
// name, mesh, cloth {elastic_stiffness, bending_stiffness, density,
// damping}, collision {enabled, self_collision, group, restitution,
// friction}, transform {translation, rotation_euler_deg, scale (number or
// 3-vector)}, pin_selection {bbox {min, max}}, angular_velocity_deg_s for
// obstacles).  As in the reference, the transform is T * Rz * Ry * Rx * S
// with the rotation values taken as radians (transform.cpp:9-17), a pin box
// selects the transformed vertices inside it (pin_selection.hpp:11-38), and
// a rotating obstacle turns about its vertex mean by dt * angular velocity
// (converted from degrees per second) before every step.  Mesh paths are
// tried as given, then relative to the scene file.
//
// --bench skips the per-step position read-back and the output, as the
// reference's --bench.  The trajectory format is described in
// silk_physics_amd/trajectory.py.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "headless/json.hpp"
#include "headless/mesh_io.hpp"
#include "silk/silk.hpp"

namespace silk::headless {
namespace {

constexpr double kPi = 3.14159265358979323846;

struct Transform {
  std::array<float, 3> t{0, 0, 0}, r{0, 0, 0}, s{1, 1, 1};
};

struct SceneObject {
  bool cloth = true;
  std::string name, mesh;
  silk::ClothConfig cc;
  silk::CollisionConfig coll;
  Transform tf;
  bool has_pin_box = false;
  std::array<float, 3> pin_min{0, 0, 0}, pin_max{0, 0, 0};
  std::array<float, 3> ang_vel_deg{0, 0, 0};
  // loaded
  std::vector<float> V;
  std::vector<int> F;
  std::vector<int> pins;
  std::array<float, 3> pivot{0, 0, 0};
  uint32_t handle = 0;
  bool moving() const { return cloth || ang_vel_deg != std::array<float, 3>{0, 0, 0}; }
};

struct Scene {
  silk::GlobalConfig g;
  int total_steps = 100;
  std::vector<SceneObject> objects;
};

float num(const Json& j, const char* what) {
  if (!j.is_number()) throw std::runtime_error(std::string("expected a number for '") + what + "'");
  return static_cast<float>(j.num);
}

std::array<float, 3> vec3(const Json& j, const char* what) {
  if (j.kind != Json::Array || j.arr.size() != 3)
    throw std::runtime_error(std::string("expected 3 numbers for '") + what + "'");
  return {num(j.arr[0], what), num(j.arr[1], what), num(j.arr[2], what)};
}

Scene parse_scene(const std::string& text) {
  const Json root = JsonParser(text).parse();
  Scene sc;
  // Reference defaults (demo/src/config.hpp:13-25).
  sc.g.dt = 0.01f;
  sc.g.max_outer_iteration = 100;
  sc.g.max_inner_iteration = 50;
  sc.g.acceleration_x = 0.0f;
  sc.g.acceleration_y = 0.0f;
  sc.g.acceleration_z = -9.8f;
  if (const Json* g = root.get("global")) {
    if (const Json* v = g->get("dt")) sc.g.dt = num(*v, "dt");
    if (const Json* v = g->get("max_outer_iteration"))
      sc.g.max_outer_iteration = static_cast<int>(num(*v, "max_outer_iteration"));
    if (const Json* v = g->get("max_inner_iteration"))
      sc.g.max_inner_iteration = static_cast<int>(num(*v, "max_inner_iteration"));
    if (const Json* v = g->get("acceleration")) {
      const auto a = vec3(*v, "acceleration");
      sc.g.acceleration_x = a[0];
      sc.g.acceleration_y = a[1];
      sc.g.acceleration_z = a[2];
    }
    if (const Json* v = g->get("total_steps")) sc.total_steps = static_cast<int>(num(*v, "total_steps"));
    if (const Json* v = g->get("linear_solver_abs_tol")) sc.g.linear_solver_abs_tol = num(*v, "tol");
    if (const Json* v = g->get("linear_solver_rel_tol_min"))
      sc.g.linear_solver_rel_tol_min = num(*v, "tol");
    if (const Json* v = g->get("linear_solver_rel_tol_max"))
      sc.g.linear_solver_rel_tol_max = num(*v, "tol");
    if (const Json* v = g->get("initial_linear_rel_tol")) sc.g.initial_linear_rel_tol = num(*v, "tol");
    if (const Json* v = g->get("admm_abs_tol")) sc.g.admm_abs_tol = num(*v, "tol");
    if (const Json* v = g->get("admm_rel_tol")) sc.g.admm_rel_tol = num(*v, "tol");
  }
  const Json* objs = root.get("objects");
  if (objs == nullptr || objs->kind != Json::Array)
    throw std::runtime_error("scene has no 'objects' list");
  for (const Json& o : objs->arr) {
    SceneObject so;
    // Collision defaults of the scene format (config.hpp:31-37).
    so.coll.is_collision_on = true;
    so.coll.is_self_collision_on = false;
    so.coll.group = 0;
    so.coll.restitution = 0.0f;
    so.coll.friction = 0.5f;
    const Json* type = o.get("type");
    if (type == nullptr || type->kind != Json::String) throw std::runtime_error("object without 'type'");
    if (type->str == "cloth") so.cloth = true;
    else if (type->str == "obstacle") so.cloth = false;
    else throw std::runtime_error("unknown object type '" + type->str + "'");
    if (const Json* v = o.get("name")) so.name = v->str;
    const Json* mesh = o.get("mesh");
    if (mesh == nullptr || mesh->kind != Json::String) throw std::runtime_error("object without 'mesh'");
    so.mesh = mesh->str;
    if (const Json* c = o.get("cloth")) {
      if (const Json* v = c->get("elastic_stiffness")) so.cc.elastic_stiffness = num(*v, "elastic_stiffness");
      if (const Json* v = c->get("bending_stiffness")) so.cc.bending_stiffness = num(*v, "bending_stiffness");
      if (const Json* v = c->get("density")) so.cc.density = num(*v, "density");
      if (const Json* v = c->get("damping")) so.cc.damping = num(*v, "damping");
    }
    if (const Json* c = o.get("collision")) {
      if (const Json* v = c->get("enabled")) so.coll.is_collision_on = v->b;
      if (const Json* v = c->get("self_collision")) so.coll.is_self_collision_on = v->b;
      if (const Json* v = c->get("group")) so.coll.group = static_cast<int>(num(*v, "group"));
      if (const Json* v = c->get("restitution")) so.coll.restitution = num(*v, "restitution");
      if (const Json* v = c->get("friction")) so.coll.friction = num(*v, "friction");
    }
    if (const Json* t = o.get("transform")) {
      if (const Json* v = t->get("translation")) so.tf.t = vec3(*v, "translation");
      if (const Json* v = t->get("rotation_euler_deg")) so.tf.r = vec3(*v, "rotation_euler_deg");
      if (const Json* v = t->get("scale")) {
        if (v->is_number()) so.tf.s = {num(*v, "scale"), num(*v, "scale"), num(*v, "scale")};
        else so.tf.s = vec3(*v, "scale");
      }
    }
    if (const Json* p = o.get("pin_selection"))
      if (const Json* b = p->get("bbox")) {
        so.has_pin_box = true;
        if (const Json* v = b->get("min")) so.pin_min = vec3(*v, "pin min");
        if (const Json* v = b->get("max")) so.pin_max = vec3(*v, "pin max");
      }
    if (const Json* v = o.get("angular_velocity_deg_s")) so.ang_vel_deg = vec3(*v, "angular_velocity_deg_s");
    sc.objects.push_back(so);
  }
  return sc;
}

// x -> T + Rz(rz) Ry(ry) Rx(rx) (S x).
void apply_transform(std::vector<float>& V, const Transform& tf) {
  const double cx = std::cos(tf.r[0]), sx = std::sin(tf.r[0]);
  const double cy = std::cos(tf.r[1]), sy = std::sin(tf.r[1]);
  const double cz = std::cos(tf.r[2]), sz = std::sin(tf.r[2]);
  // R = Rz * Ry * Rx.
  const double R[3][3] = {
      {cz * cy, cz * sy * sx - sz * cx, cz * sy * cx + sz * sx},
      {sz * cy, sz * sy * sx + cz * cx, sz * sy * cx - cz * sx},
      {-sy, cy * sx, cy * cx}};
  for (size_t i = 0; i < V.size(); i += 3) {
    const double p[3] = {tf.s[0] * V[i], tf.s[1] * V[i + 1], tf.s[2] * V[i + 2]};
    for (int r = 0; r < 3; ++r)
      V[i + r] = static_cast<float>(R[r][0] * p[0] + R[r][1] * p[1] + R[r][2] * p[2] + tf.t[r]);
  }
}

std::string resolve(const std::string& mesh, const std::filesystem::path& scene_dir) {
  if (std::filesystem::exists(mesh)) return mesh;
  const auto p = scene_dir / mesh;
  if (std::filesystem::exists(p)) return p.string();
  return mesh;
}

void check(const silk::Result& r, const std::string& what) {
  if (!r) throw std::runtime_error(what + ": " + r.to_string());
}

template <typename T>
void put(std::ofstream& f, const T& v) {
  f.write(reinterpret_cast<const char*>(&v), sizeof(T));
}

int run(int argc, char** argv) {
  std::string config, output;
  bool bench = false, quiet = false, check_only = false;
  int steps_override = -1;
  silk::Backend backend = silk::Backend::GPU;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) throw std::runtime_error("missing value after " + a);
      return argv[++i];
    };
    if (a == "--config") config = next();
    else if (a == "--output") output = next();
    else if (a == "--bench") bench = true;
    else if (a == "--quiet") quiet = true;
    else if (a == "--check") check_only = true;
    else if (a == "--steps") steps_override = std::stoi(next());
    else if (a == "--backend") {
      const std::string b = next();
      if (b == "gpu") backend = silk::Backend::GPU;
      else if (b == "cpu") backend = silk::Backend::CPU;
      else throw std::runtime_error("unknown backend '" + b + "'");
    } else {
      throw std::runtime_error("unknown argument '" + a + "'");
    }
  }
  if (config.empty()) throw std::runtime_error("--config is required");
  std::ifstream cf(config);
  if (!cf) throw std::runtime_error("cannot open scene file '" + config + "'");
  std::stringstream ss;
  ss << cf.rdbuf();
  Scene sc = parse_scene(ss.str());
  if (steps_override >= 0) sc.total_steps = steps_override;
  const auto scene_dir = std::filesystem::path(config).parent_path();

  for (auto& o : sc.objects) {
    TriMesh m = read_mesh(resolve(o.mesh, scene_dir));
    if (m.V.empty() || m.F.empty()) throw std::runtime_error("empty mesh '" + o.mesh + "'");
    apply_transform(m.V, o.tf);
    o.V = std::move(m.V);
    o.F = std::move(m.F);
    const int nv = static_cast<int>(o.V.size() / 3);
    if (o.has_pin_box) {
      std::array<float, 3> lo = o.pin_min, hi = o.pin_max;
      for (int d = 0; d < 3; ++d)
        if (lo[d] > hi[d]) std::swap(lo[d], hi[d]);
      for (int i = 0; i < nv; ++i) {
        bool in = true;
        for (int d = 0; d < 3; ++d) in = in && o.V[3 * i + d] >= lo[d] && o.V[3 * i + d] <= hi[d];
        if (in) o.pins.push_back(i);
      }
    }
    for (int i = 0; i < nv; ++i)
      for (int d = 0; d < 3; ++d) o.pivot[d] += o.V[3 * i + d] / static_cast<float>(nv);
  }
  if (check_only) {
    // One JSON line per object: what the scene resolves to, no device work.
    for (const auto& o : sc.objects) {
      float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
      for (size_t i = 0; i < o.V.size(); ++i) {
        lo[i % 3] = std::min(lo[i % 3], o.V[i]);
        hi[i % 3] = std::max(hi[i % 3], o.V[i]);
      }
      std::printf("{\"name\": \"%s\", \"cloth\": %s, \"vertices\": %zu, \"triangles\": %zu, "
                  "\"pins\": %zu, \"group\": %d, \"bbox\": [%.6g, %.6g, %.6g, %.6g, %.6g, %.6g]}\n",
                  o.name.c_str(), o.cloth ? "true" : "false", o.V.size() / 3, o.F.size() / 3,
                  o.pins.size(), o.coll.group, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
    }
    std::printf("{\"steps\": %d, \"dt\": %.6g, \"max_outer_iteration\": %d, "
                "\"max_inner_iteration\": %d}\n",
                sc.total_steps, sc.g.dt, sc.g.max_outer_iteration, sc.g.max_inner_iteration);
    return 0;
  }
  silk::World world;
  check(world.init(backend), "init");
  for (auto& o : sc.objects) {
    silk::MeshConfig mc{std::span<const float>(o.V), std::span<const int>(o.F)};
    if (o.cloth)
      check(world.add_cloth(o.cc, o.coll, mc, std::span<const int>(o.pins), o.handle),
            "add_cloth '" + o.name + "'");
    else
      check(world.add_obstacle(o.coll, mc, o.handle), "add_obstacle '" + o.name + "'");
  }
  check(world.set_global_config(sc.g), "set_global_config");

  std::ofstream out;
  std::streampos nframes_pos;
  uint32_t nframes = 0;
  auto write_frame = [&](float time) {
    put(out, time);
    for (const auto& o : sc.objects)
      if (o.moving()) out.write(reinterpret_cast<const char*>(o.V.data()), 4 * o.V.size());
    ++nframes;
  };
  if (!bench && !output.empty()) {
    out.open(output, std::ios::binary);
    if (!out) throw std::runtime_error("cannot write '" + output + "'");
    out.write("SILKTRJ1", 8);
    put(out, static_cast<uint32_t>(sc.objects.size()));
    for (const auto& o : sc.objects) {
      put(out, static_cast<uint32_t>(o.cloth ? 0 : 1));
      put(out, static_cast<uint32_t>(o.moving() ? 1 : 0));
      put(out, static_cast<uint32_t>(o.name.size()));
      out.write(o.name.data(), static_cast<std::streamsize>(o.name.size()));
      put(out, static_cast<uint32_t>(o.V.size() / 3));
      put(out, static_cast<uint32_t>(o.F.size() / 3));
      out.write(reinterpret_cast<const char*>(o.V.data()), 4 * o.V.size());
      out.write(reinterpret_cast<const char*>(o.F.data()), 4 * o.F.size());
    }
    nframes_pos = out.tellp();
    put(out, nframes);
    write_frame(0.0f);
  }

  const auto t0 = std::chrono::steady_clock::now();
  for (int step = 0; step < sc.total_steps; ++step) {
    for (auto& o : sc.objects) {
      if (o.cloth || !o.moving()) continue;
      Transform rot;
      for (int d = 0; d < 3; ++d)
        rot.r[d] = static_cast<float>(sc.g.dt * o.ang_vel_deg[d] * kPi / 180.0);
      for (size_t i = 0; i < o.V.size(); i += 3)
        for (int d = 0; d < 3; ++d) o.V[i + d] -= o.pivot[d];
      apply_transform(o.V, rot);
      for (size_t i = 0; i < o.V.size(); i += 3)
        for (int d = 0; d < 3; ++d) o.V[i + d] += o.pivot[d];
      check(world.set_obstacle_position(o.handle, std::span<const float>(o.V)),
            "set_obstacle_position '" + o.name + "'");
    }
    check(world.solver_step(), "solver step " + std::to_string(step));
    const float time = static_cast<float>((step + 1) * static_cast<double>(sc.g.dt));
    if (!quiet) std::printf("step %d time %.6g\n", step, time);
    if (!bench) {
      for (auto& o : sc.objects)
        if (o.cloth) check(world.get_cloth_position(o.handle, std::span<float>(o.V)),
                           "get_cloth_position '" + o.name + "'");
      if (out.is_open()) write_frame(time);
    }
  }
  const double wall =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const bool wrote = out.is_open();
  if (wrote) {
    out.seekp(nframes_pos);
    put(out, nframes);
    out.close();
  }
  std::printf("done: %d steps in %.3f s (%.3f steps/s)\n", sc.total_steps, wall,
              sc.total_steps / std::max(wall, 1e-9));
  if (wrote) std::printf("trajectory: %s (%u frames)\n", output.c_str(), nframes);
  return 0;
}

}  // namespace
}  // namespace silk::headless

int main(int argc, char** argv) {
  try {
    return silk::headless::run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "silk_headless: %s\n", e.what());
    return 1;
  }
}
