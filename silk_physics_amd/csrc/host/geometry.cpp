// Host-side mesh validation, static cloth operators, partitioner and MAS
// level hierarchy.  See geometry.hpp for the reference anchors.
#include "host/geometry.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <functional>
#include <numeric>
#include <unordered_map>

namespace silk::amd {

namespace {

using V3 = std::array<double, 3>;

V3 vload(const std::vector<float>& V, int i) {
  return {V[3 * i], V[3 * i + 1], V[3 * i + 2]};
}
V3 sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
V3 cross(const V3& a, const V3& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
          a[0] * b[1] - a[1] * b[0]};
}
double norm(const V3& a) { return std::sqrt(dot(a, a)); }

bool check_schema(const float* verts, int64_t nvf, const int* faces,
                  int64_t nff, int min_v, int min_f, std::string& why) {
  if (verts == nullptr || (faces == nullptr && nff != 0) || nvf < 0 ||
      nff < 0) {
    why = "null vertex or face buffer";
    return false;
  }
  if (nvf % 3 != 0 || nff % 3 != 0) {
    why = "vertex and face arrays must hold multiples of 3 values";
    return false;
  }
  if (nvf / 3 < min_v) {
    why = "too few vertices";
    return false;
  }
  if (nff / 3 < min_f) {
    why = "too few faces";
    return false;
  }
  if (nvf / 3 > (int64_t{1} << 28) || nff / 3 > (int64_t{1} << 28)) {
    why = "mesh too large";
    return false;
  }
  return true;
}

bool check_basic(TriMesh& m, const float* verts, int64_t nvf, const int* faces,
                 int64_t nff, std::string& why) {
  m.V.assign(verts, verts + nvf);
  m.F.assign(faces, faces + nff);
  for (float v : m.V) {
    if (!std::isfinite(v)) {
      why = "vertex positions contain NaN/Inf";
      return false;
    }
  }
  const int nv = m.nv();
  for (int f = 0; f < m.nf(); ++f) {
    int a = m.F[3 * f], b = m.F[3 * f + 1], c = m.F[3 * f + 2];
    if (a < 0 || b < 0 || c < 0 || a >= nv || b >= nv || c >= nv) {
      why = "face index out of range";
      return false;
    }
    if (a == b || b == c || a == c) {
      why = "triangle with repeated vertex index";
      return false;
    }
  }
  return true;
}

// Unique undirected edges as sorted (min, max) pairs, and the float average
// edge length accumulated in that order.
void build_edges(TriMesh& m) {
  std::vector<std::pair<int, int>> e;
  e.reserve(3 * m.nf());
  for (int f = 0; f < m.nf(); ++f) {
    for (int k = 0; k < 3; ++k) {
      int a = m.F[3 * f + k], b = m.F[3 * f + (k + 1) % 3];
      e.emplace_back(std::min(a, b), std::max(a, b));
    }
  }
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  m.E.resize(2 * e.size());
  float sum = 0.0f;
  for (size_t i = 0; i < e.size(); ++i) {
    m.E[2 * i] = e[i].first;
    m.E[2 * i + 1] = e[i].second;
    float dx = m.V[3 * e[i].first] - m.V[3 * e[i].second];
    float dy = m.V[3 * e[i].first + 1] - m.V[3 * e[i].second + 1];
    float dz = m.V[3 * e[i].first + 2] - m.V[3 * e[i].second + 2];
    sum += std::sqrt(dx * dx + dy * dy + dz * dz);
  }
  m.avg_edge_length = e.empty() ? 0.0f : sum / static_cast<float>(e.size());
}

struct Dsu {
  std::vector<int> p;
  explicit Dsu(int n) : p(n) { std::iota(p.begin(), p.end(), 0); }
  int find(int x) {
    while (p[x] != x) {
      p[x] = p[p[x]];
      x = p[x];
    }
    return x;
  }
  void unite(int a, int b) {
    a = find(a);
    b = find(b);
    if (a != b) p[std::max(a, b)] = std::min(a, b);
  }
};

}  // namespace

bool make_obstacle_mesh(const float* verts, int64_t nvf, const int* faces,
                        int64_t nff, TriMesh& out, std::string& why) {
  if (!check_schema(verts, nvf, faces, nff, 1, 0, why)) return false;
  TriMesh m;
  if (!check_basic(m, verts, nvf, faces, nff, why)) return false;
  build_edges(m);
  out = std::move(m);
  return true;
}

bool make_cloth_mesh(const float* verts, int64_t nvf, const int* faces,
                     int64_t nff, TriMesh& out, std::string& why) {
  if (!check_schema(verts, nvf, faces, nff, 3, 1, why)) return false;
  TriMesh m;
  if (!check_basic(m, verts, nvf, faces, nff, why)) return false;
  const int nv = m.nv(), nf = m.nf();

  // Internal angles within [0.5, 179.5] degrees (mesh.cpp:156-186).
  const double min_rad = 0.5 * M_PI / 180.0;
  const double max_rad = M_PI - min_rad;
  for (int f = 0; f < nf; ++f) {
    for (int k = 0; k < 3; ++k) {
      V3 p = vload(m.V, m.F[3 * f + k]);
      V3 a = sub(vload(m.V, m.F[3 * f + (k + 1) % 3]), p);
      V3 b = sub(vload(m.V, m.F[3 * f + (k + 2) % 3]), p);
      double la = norm(a), lb = norm(b);
      double ang = (la == 0.0 || lb == 0.0)
                       ? 0.0
                       : std::acos(std::clamp(dot(a, b) / (la * lb), -1.0, 1.0));
      if (!(ang >= min_rad && ang <= max_rad)) {
        why = "poor triangle angle quality";
        return false;
      }
    }
  }

  // Edge manifold: at most two faces per undirected edge.
  std::vector<std::array<int, 3>> he;  // (min, max, face)
  he.reserve(3 * nf);
  for (int f = 0; f < nf; ++f) {
    for (int k = 0; k < 3; ++k) {
      int a = m.F[3 * f + k], b = m.F[3 * f + (k + 1) % 3];
      he.push_back({std::min(a, b), std::max(a, b), f});
    }
  }
  std::sort(he.begin(), he.end());
  Dsu faces_dsu(nf);
  for (size_t i = 0; i < he.size();) {
    size_t j = i;
    while (j < he.size() && he[j][0] == he[i][0] && he[j][1] == he[i][1]) ++j;
    if (j - i > 2) {
      why = "non-manifold edge";
      return false;
    }
    for (size_t k = i + 1; k < j; ++k) faces_dsu.unite(he[i][2], he[k][2]);
    i = j;
  }

  // Vertex manifold: the faces around each vertex form one fan, connected
  // through edges that contain the vertex (mesh.cpp:202-214).
  std::vector<std::vector<int>> vf(nv);
  for (int f = 0; f < nf; ++f)
    for (int k = 0; k < 3; ++k) vf[m.F[3 * f + k]].push_back(f);
  for (int v = 0; v < nv; ++v) {
    const auto& fs = vf[v];
    if (fs.size() <= 1) continue;
    // Map each incident face to the two other corners; union faces sharing a
    // corner (i.e. an edge through v).
    std::unordered_map<int, int> first_face_of_corner;
    Dsu d(static_cast<int>(fs.size()));
    for (int i = 0; i < static_cast<int>(fs.size()); ++i) {
      int f = fs[i];
      for (int k = 0; k < 3; ++k) {
        int c = m.F[3 * f + k];
        if (c == v) continue;
        auto it = first_face_of_corner.find(c);
        if (it == first_face_of_corner.end())
          first_face_of_corner.emplace(c, i);
        else
          d.unite(it->second, i);
      }
    }
    int root = d.find(0);
    for (int i = 1; i < static_cast<int>(fs.size()); ++i) {
      if (d.find(i) != root) {
        why = "non-manifold vertex";
        return false;
      }
    }
  }

  // Single edge-connected component (mesh.cpp:220-229).
  int root = faces_dsu.find(0);
  for (int f = 1; f < nf; ++f) {
    if (faces_dsu.find(f) != root) {
      why = "mesh has more than one connected component";
      return false;
    }
  }

  build_edges(m);
  out = std::move(m);
  return true;
}

ClothGeometry build_cloth_geometry(const TriMesh& mesh) {
  const int nv = mesh.nv(), nf = mesh.nf();
  ClothGeometry g;
  g.mass.assign(nv, 0.0);
  g.area.assign(nf, 0.0);
  g.face_param.assign(4 * nf, 0.0f);

  // COO accumulators for the two vertex-scalar operators.
  std::vector<std::tuple<int, int, double>> elastic_coo;
  elastic_coo.reserve(9 * nf);
  std::vector<std::tuple<int, int, double>> cot_coo;
  cot_coo.reserve(12 * nf);

  for (int f = 0; f < nf; ++f) {
    const int id[3] = {mesh.F[3 * f], mesh.F[3 * f + 1], mesh.F[3 * f + 2]};
    const V3 p[3] = {vload(mesh.V, id[0]), vload(mesh.V, id[1]),
                     vload(mesh.V, id[2])};
    V3 e0 = sub(p[1], p[0]);
    V3 e1 = sub(p[2], p[0]);
    V3 n = cross(e0, e1);
    double dbl_area = norm(n);
    double area = 0.5 * dbl_area;
    g.area[f] = area;

    // Local 2-D chart with e0 as x axis (cloth_assembly_l2_cache.cpp:27-69).
    double a = norm(e0);
    V3 by = cross(n, e0);
    double lby = norm(by);
    double b = dot(e0, e1) / a;
    double c = dot(by, e1) / lby;
    double ia = 1.0 / a, ib = -b / (a * c), ic = 1.0 / c;
    g.face_param[4 * f + 0] = static_cast<float>(ia);
    g.face_param[4 * f + 1] = static_cast<float>(ib);
    g.face_param[4 * f + 2] = static_cast<float>(ic);
    g.face_param[4 * f + 3] = static_cast<float>(std::sqrt(area));

    // B rows over the three corners.
    const double B[2][3] = {{-ia, ia, 0.0}, {-(ib + ic), ib, ic}};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        elastic_coo.emplace_back(
            id[i], id[j], area * (B[0][i] * B[0][j] + B[1][i] * B[1][j]));

    // Half-cotangent weights and mixed Voronoi areas (Meyer et al. 2003).
    double cot[3];
    double cosv[3];
    double l2[3];  // squared length of the edge opposite corner k
    for (int k = 0; k < 3; ++k) {
      V3 u = sub(p[(k + 1) % 3], p[k]);
      V3 w = sub(p[(k + 2) % 3], p[k]);
      cot[k] = dot(u, w) / dbl_area;
      cosv[k] = dot(u, w);
      V3 opp = sub(p[(k + 2) % 3], p[(k + 1) % 3]);
      l2[k] = dot(opp, opp);
    }
    for (int k = 0; k < 3; ++k) {
      int i = id[(k + 1) % 3], j = id[(k + 2) % 3];
      double wgt = 0.5 * cot[k];
      cot_coo.emplace_back(i, j, wgt);
      cot_coo.emplace_back(j, i, wgt);
      cot_coo.emplace_back(i, i, -wgt);
      cot_coo.emplace_back(j, j, -wgt);
    }
    int obtuse = -1;
    for (int k = 0; k < 3; ++k)
      if (cosv[k] < 0.0) obtuse = k;
    for (int k = 0; k < 3; ++k) {
      double mk;
      if (obtuse < 0) {
        // corner k touches the two edges adjacent to it: the edge opposite
        // corner k+1 (length l2[k+1]) weighted by cot at k+1, and likewise k+2.
        mk = 0.125 * (l2[(k + 1) % 3] * cot[(k + 1) % 3] +
                      l2[(k + 2) % 3] * cot[(k + 2) % 3]);
      } else {
        mk = (obtuse == k) ? 0.5 * area : 0.25 * area;
      }
      g.mass[id[k]] += mk;
    }
  }

  auto to_csr = [nv](std::vector<std::tuple<int, int, double>>& coo) {
    std::sort(coo.begin(), coo.end(), [](const auto& x, const auto& y) {
      return std::get<0>(x) != std::get<0>(y) ? std::get<0>(x) < std::get<0>(y)
                                              : std::get<1>(x) < std::get<1>(y);
    });
    CsrMatrix m;
    m.n = nv;
    m.row_ptr.assign(nv + 1, 0);
    for (size_t i = 0; i < coo.size();) {
      size_t j = i;
      double s = 0.0;
      while (j < coo.size() && std::get<0>(coo[j]) == std::get<0>(coo[i]) &&
             std::get<1>(coo[j]) == std::get<1>(coo[i])) {
        s += std::get<2>(coo[j]);
        ++j;
      }
      m.col.push_back(std::get<1>(coo[i]));
      m.val.push_back(s);
      m.row_ptr[std::get<0>(coo[i]) + 1]++;
      i = j;
    }
    for (int r = 0; r < nv; ++r) m.row_ptr[r + 1] += m.row_ptr[r];
    return m;
  };

  g.elastic = to_csr(elastic_coo);
  CsrMatrix C = to_csr(cot_coo);

  // CWC = C^T W C with W = diag(1/mass); C is symmetric.
  std::vector<std::tuple<int, int, double>> cwc_coo;
  for (int k = 0; k < nv; ++k) {
    double wk = 1.0 / g.mass[k];
    for (int a = C.row_ptr[k]; a < C.row_ptr[k + 1]; ++a)
      for (int b = C.row_ptr[k]; b < C.row_ptr[k + 1]; ++b)
        cwc_coo.emplace_back(C.col[a], C.col[b], C.val[a] * C.val[b] * wk);
  }
  g.bending = to_csr(cwc_coo);

  g.bending_rest.assign(3 * nv, 0.0);
  for (int r = 0; r < nv; ++r)
    for (int a = g.bending.row_ptr[r]; a < g.bending.row_ptr[r + 1]; ++a)
      for (int d = 0; d < 3; ++d)
        g.bending_rest[3 * r + d] +=
            g.bending.val[a] * mesh.V[3 * g.bending.col[a] + d];
  return g;
}

CsrMatrix combine_operator(const ClothGeometry& g, double k, double kb) {
  const CsrMatrix& A = g.elastic;
  const CsrMatrix& B = g.bending;
  CsrMatrix m;
  m.n = A.n;
  m.row_ptr.assign(A.n + 1, 0);
  for (int r = 0; r < A.n; ++r) {
    int ia = A.row_ptr[r], ea = A.row_ptr[r + 1];
    int ib = B.row_ptr[r], eb = B.row_ptr[r + 1];
    while (ia < ea || ib < eb) {
      int ca = ia < ea ? A.col[ia] : INT32_MAX;
      int cb = ib < eb ? B.col[ib] : INT32_MAX;
      int c = std::min(ca, cb);
      double v = 0.0;
      if (ca == c) v += k * A.val[ia++];
      if (cb == c) v += kb * B.val[ib++];
      m.col.push_back(c);
      m.val.push_back(v);
    }
    m.row_ptr[r + 1] = static_cast<int>(m.col.size());
  }
  return m;
}

Graph pattern_graph(const CsrMatrix& m) {
  Graph g;
  g.n = m.n;
  g.row_ptr.assign(m.n + 1, 0);
  for (int r = 0; r < m.n; ++r) {
    for (int a = m.row_ptr[r]; a < m.row_ptr[r + 1]; ++a)
      if (m.col[a] != r) g.col.push_back(m.col[a]);
    g.row_ptr[r + 1] = static_cast<int>(g.col.size());
  }
  return g;
}

Partition partition_vertices(const std::vector<float>& V, int max_part_size) {
  const int n = static_cast<int>(V.size() / 3);
  Partition p;
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::vector<int> part_of(n, 0);
  int nparts = 0;
  if (n <= max_part_size) {
    nparts = 1;
  } else {
    const int target = std::max(1, max_part_size - 2);
    const int k_total = (n + target - 1) / target;
    // Recursive coordinate bisection: split the index range [lo, hi) into k
    // parts along its widest axis with proportional counts.
    std::function<void(int, int, int)> rcb = [&](int lo, int hi, int k) {
      if (k == 1) {
        for (int i = lo; i < hi; ++i) part_of[order[i]] = nparts;
        ++nparts;
        return;
      }
      double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
      for (int i = lo; i < hi; ++i)
        for (int d = 0; d < 3; ++d) {
          mn[d] = std::min(mn[d], static_cast<double>(V[3 * order[i] + d]));
          mx[d] = std::max(mx[d], static_cast<double>(V[3 * order[i] + d]));
        }
      int axis = 0;
      for (int d = 1; d < 3; ++d)
        if (mx[d] - mn[d] > mx[axis] - mn[axis]) axis = d;
      std::sort(order.begin() + lo, order.begin() + hi, [&](int a, int b) {
        float va = V[3 * a + axis], vb = V[3 * b + axis];
        return va != vb ? va < vb : a < b;
      });
      const int cnt = hi - lo;
      const int k1 = k / 2;
      int64_t n1 = (2 * int64_t{cnt} * k1 + k) / (2 * int64_t{k});
      n1 = std::clamp<int64_t>(n1, k1, cnt - (k - k1));
      rcb(lo, lo + static_cast<int>(n1), k1);
      rcb(lo + static_cast<int>(n1), hi, k - k1);
    };
    rcb(0, n, k_total);
  }
  p.offsets.assign(nparts + 1, 0);
  for (int v = 0; v < n; ++v) p.offsets[part_of[v] + 1]++;
  for (int i = 0; i < nparts; ++i) p.offsets[i + 1] += p.offsets[i];
  std::vector<int> cursor(p.offsets.begin(), p.offsets.end() - 1);
  p.perm.assign(n, 0);
  p.inv_perm.assign(n, 0);
  for (int v = 0; v < n; ++v) {
    p.perm[v] = cursor[part_of[v]]++;
    p.inv_perm[p.perm[v]] = v;
  }
  return p;
}

TriMesh permute_mesh(const TriMesh& mesh, const Partition& part) {
  TriMesh m = mesh;
  for (int nv = 0; nv < mesh.nv(); ++nv) {
    int old = part.inv_perm[nv];
    for (int d = 0; d < 3; ++d) m.V[3 * nv + d] = mesh.V[3 * old + d];
  }
  for (size_t i = 0; i < m.F.size(); ++i) m.F[i] = part.perm[mesh.F[i]];
  for (size_t i = 0; i < m.E.size(); ++i) m.E[i] = part.perm[mesh.E[i]];
  return m;
}

MasTopology build_mas_topology(const Graph& g,
                               const std::vector<int>& part_offsets) {
  constexpr int B = MasTopology::kBank;
  MasTopology t;
  t.n = g.n;
  t.nbanks0 = static_cast<int>(part_offsets.size()) - 1;
  t.pad_of.assign(g.n, -1);
  for (int b = 0; b < t.nbanks0; ++b)
    for (int v = part_offsets[b]; v < part_offsets[b + 1]; ++v)
      t.pad_of[v] = b * B + (v - part_offsets[b]);
  t.level_nodes.push_back(t.nbanks0 * B);
  t.level_banks.push_back(t.nbanks0);

  // Components inside 32-slot banks; component ids ordered by bank and then
  // by lowest member slot (mas_preconditioner.cu:162-237,298-374).
  auto label = [&](int nnodes, const std::function<int(int)>& slot_of_node,
                   Dsu& d, std::vector<int>& out_id) -> int {
    // root -> lowest slot; ids assigned bank by bank in lowest-slot order.
    const int nb = (nnodes + B - 1) / B;
    std::vector<int> comp(nnodes);
    for (int i = 0; i < nnodes; ++i) comp[i] = d.find(i);
    out_id.assign(nnodes, -1);
    int next = 0;
    for (int bk = 0; bk < nb; ++bk) {
      // nodes of this bank in slot order
      std::vector<std::pair<int, int>> members;  // (slot, node)
      for (int i = bk * B; i < std::min(nnodes, (bk + 1) * B); ++i)
        members.emplace_back(slot_of_node(i), i);
      std::sort(members.begin(), members.end());
      std::unordered_map<int, int> root_id;
      for (auto& [slot, node] : members) {
        (void)slot;
        int r = comp[node];
        auto it = root_id.find(r);
        if (it == root_id.end()) it = root_id.emplace(r, next++).first;
        out_id[node] = it->second;
      }
    }
    return next;
  };

  // Level 0: nodes are padded slots (only real ones participate).
  const int np = t.nbanks0 * B;
  std::vector<int> real_of_pad(np, -1);
  for (int v = 0; v < g.n; ++v) real_of_pad[t.pad_of[v]] = v;
  {
    Dsu d(np);
    for (int v = 0; v < g.n; ++v)
      for (int a = g.row_ptr[v]; a < g.row_ptr[v + 1]; ++a) {
        int u = g.col[a];
        if (t.pad_of[v] / B == t.pad_of[u] / B) d.unite(t.pad_of[v], t.pad_of[u]);
      }
    // Label only real slots: compact ids over real nodes.
    std::vector<int> pid;
    const int nb = t.nbanks0;
    std::vector<int> comp(np);
    for (int i = 0; i < np; ++i) comp[i] = d.find(i);
    std::vector<int> id0(g.n, -1);
    int next = 0;
    for (int bk = 0; bk < nb; ++bk) {
      std::unordered_map<int, int> root_id;
      for (int s = 0; s < B; ++s) {
        int pi = bk * B + s;
        if (real_of_pad[pi] < 0) continue;
        int r = comp[pi];
        auto it = root_id.find(r);
        if (it == root_id.end()) it = root_id.emplace(r, next++).first;
        id0[real_of_pad[pi]] = it->second;
      }
    }
    t.cmap.push_back(id0);
    t.level_nodes.push_back(next);
    t.level_banks.push_back((next + B - 1) / B);
    t.coarse_levels = 1;
  }

  for (int lv = 1; lv < MasTopology::kMaxCoarse; ++lv) {
    const std::vector<int>& prev = t.cmap.back();
    const int nprev = t.level_nodes.back();
    Dsu d(nprev);
    for (int v = 0; v < g.n; ++v)
      for (int a = g.row_ptr[v]; a < g.row_ptr[v + 1]; ++a) {
        int I = prev[v], J = prev[g.col[a]];
        if (I != J && I / B == J / B) d.unite(I, J);
      }
    std::vector<int> node_id;
    int next = label(nprev, [](int i) { return i % B; }, d, node_id);
    if (next == nprev) break;
    std::vector<int> idl(g.n);
    for (int v = 0; v < g.n; ++v) idl[v] = node_id[prev[v]];
    t.cmap.push_back(idl);
    t.level_nodes.push_back(next);
    t.level_banks.push_back((next + B - 1) / B);
    t.coarse_levels = lv + 1;
  }
  return t;
}

std::vector<double> mas_static_blocks(const MasTopology& topo,
                                      const CsrMatrix& K) {
  constexpr int B = MasTopology::kBank;
  int total_banks = 0;
  for (int b : topo.level_banks) total_banks += b;
  std::vector<double> out(static_cast<size_t>(total_banks) * B * B, 0.0);
  // Level 0.
  for (int i = 0; i < K.n; ++i) {
    int pi = topo.pad_of[i];
    for (int a = K.row_ptr[i]; a < K.row_ptr[i + 1]; ++a) {
      int pj = topo.pad_of[K.col[a]];
      if (pi / B != pj / B) continue;
      out[static_cast<size_t>(pi / B) * B * B + (pi % B) * B + (pj % B)] +=
          K.val[a];
    }
  }
  size_t bank_base = topo.level_banks[0];
  for (int l = 0; l < topo.coarse_levels; ++l) {
    const auto& map = topo.cmap[l];
    for (int i = 0; i < K.n; ++i) {
      int I = map[i];
      for (int a = K.row_ptr[i]; a < K.row_ptr[i + 1]; ++a) {
        int J = map[K.col[a]];
        if (I / B != J / B) continue;
        out[(bank_base + I / B) * B * B + (I % B) * B + (J % B)] += K.val[a];
      }
    }
    bank_base += topo.level_banks[l + 1];
  }
  return out;
}

}  // namespace silk::amd
