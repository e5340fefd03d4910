// Host-side mesh validation and static cloth operators.
//
// Everything here runs once per cloth (or once per dt change) on the CPU in
// double precision and is uploaded to the device as float.  The reference
// builds the same quantities with libigl/Eigen in float
// (silk/src/common/mesh.cpp:186-266, cloth_assembly_l2_cache.cpp:27-137,
// backend/cuda/assembly/cloth_assembly_l1_cache.cu:16-97).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace silk::amd {

/// Validated triangle mesh: V (3*nv), F (3*nf), E (2*ne, sorted unique
/// (min,max) pairs).
struct TriMesh {
  std::vector<float> V;
  std::vector<int> F;
  std::vector<int> E;
  float avg_edge_length = 0.0f;
  int nv() const { return static_cast<int>(V.size() / 3); }
  int nf() const { return static_cast<int>(F.size() / 3); }
  int ne() const { return static_cast<int>(E.size() / 2); }
};

/// Cloth mesh checks of silk/src/common/mesh.cpp:186-231: schema, finite
/// positions, index range, no repeated corner, every internal angle in
/// [0.5 deg, 179.5 deg], edge manifold, vertex manifold, one edge-connected
/// component.  Returns false with a reason on failure.
bool make_cloth_mesh(const float* verts, int64_t nverts_floats,
                     const int* faces, int64_t nfaces_ints, TriMesh& out,
                     std::string& why);

/// Obstacle mesh checks of silk/src/common/mesh.cpp:233-266 (schema, finite,
/// index range, no repeated corner).
bool make_obstacle_mesh(const float* verts, int64_t nverts_floats,
                        const int* faces, int64_t nfaces_ints, TriMesh& out,
                        std::string& why);

/// Symmetric sparse matrix in CSR form with sorted columns (diagonal
/// included).  Used for the vertex-scalar part of the PD operator.
struct CsrMatrix {
  int n = 0;
  std::vector<int> row_ptr;
  std::vector<int> col;
  std::vector<double> val;
};

/// Static per-mesh quantities (reference ClothAssemblyL2Cache,
/// cloth_assembly_l2_cache.hpp:22-61), for a mesh already in solver order.
///
/// The reference's 6x9 per-face Jacobian is kron(B, I3) with
/// B = (D dX^-1)^T (cloth_assembly_l2_cache.cpp:27-69).  dX is upper
/// triangular, so B is fixed by three numbers per face:
///   F(:,0) = e0 * ia,   F(:,1) = e0 * ib + e1 * ic,
/// with e0 = x1 - x0, e1 = x2 - x0.
struct ClothGeometry {
  std::vector<double> mass;       // Voronoi area per vertex (no density)
  std::vector<double> area;       // per face
  std::vector<float> face_param;  // per face: ia, ib, ic, sqrt(area)
  CsrMatrix elastic;              // sum_f area_f B_f^T B_f (vertex scalar)
  CsrMatrix bending;              // C^T diag(1/mass) C     (vertex scalar)
  std::vector<double> bending_rest;  // bending * V (3*nv, interleaved xyz)
};

ClothGeometry build_cloth_geometry(const TriMesh& mesh);

/// k * elastic + kb * bending, on the union pattern.
CsrMatrix combine_operator(const ClothGeometry& g, double k, double kb);

/// Vertex graph of a CSR pattern without self loops.
struct Graph {
  int n = 0;
  std::vector<int> row_ptr;
  std::vector<int> col;
};
Graph pattern_graph(const CsrMatrix& m);

/// Partition of the mesh vertices into parts of at most 32 vertices
/// (reference uses KaMinPar with k = ceil(n/30), graph_partition.cpp:14-92;
/// this one is recursive coordinate bisection on the rest positions).
/// perm: original vertex -> solver order; parts are contiguous in solver order
/// and vertices keep their original relative order inside a part
/// (mesh_partition.cu:56-105).
struct Partition {
  std::vector<int> perm;
  std::vector<int> inv_perm;
  std::vector<int> offsets;  // nparts + 1
};
Partition partition_vertices(const std::vector<float>& V, int max_part_size);

/// Relabel a mesh by a partition permutation (cloth_assembler.cu:30-48).
TriMesh permute_mesh(const TriMesh& mesh, const Partition& part);

/// MAS level hierarchy for 32-slot banks (reference build_padded_topology and
/// build_coarse_space, mas_preconditioner.cu:99-478).
struct MasTopology {
  static constexpr int kBank = 32;
  static constexpr int kMaxCoarse = 4;
  int n = 0;                   // real nodes
  int nbanks0 = 0;             // level-0 banks (= parts)
  std::vector<int> pad_of;     // real -> padded id (bank*32 + slot)
  int coarse_levels = 0;       // number of coarse levels used (0..4)
  std::vector<int> level_nodes;        // node count per level (L0 = padded)
  std::vector<int> level_banks;        // bank count per level
  std::vector<std::vector<int>> cmap;  // [l][real] -> node id at coarse level l
};
MasTopology build_mas_topology(const Graph& g,
                               const std::vector<int>& part_offsets);

/// Static part of every MAS bank matrix (row-major 32x32 per bank, all levels
/// concatenated level by level): level 0 holds K restricted to each bank, a
/// coarse level holds the Galerkin sum P^T K P restricted to each bank.
std::vector<double> mas_static_blocks(const MasTopology& topo,
                                      const CsrMatrix& K);

}  // namespace silk::amd
