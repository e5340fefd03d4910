// Triangle mesh readers for scene files: Wavefront OBJ (v / f, polygons fanned
// into triangles, v/vt/vn references and negative indices accepted), OFF,
// PLY (ascii and binary little endian; float or double vertex coordinates,
// any list type for faces) and STL (ascii and binary; vertices merged by
// exact position).  The reference reads the same formats through libigl's
// read_triangle_mesh (demo/src/gui_utils.cpp).
#pragma once

#include <string>
#include <vector>

namespace silk::headless {

struct TriMesh {
  std::vector<float> V;  // 3 per vertex
  std::vector<int> F;    // 3 per triangle
};

/// Reads a mesh by file extension (.obj, .off, .ply, .stl, case-insensitive).
/// Throws std::runtime_error with the reason on failure.
TriMesh read_mesh(const std::string& path);

}  // namespace silk::headless
