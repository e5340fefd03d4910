// Collision subsystem of the step on the device.
//
// Reference pipeline (silk/src/backend/cuda/main_loop.cu:186-277):
//   frame start   refit colliders at t0 = t1 = x, DCD, contact set
//   every segment refit swept colliders x -> x_pred, CCD, earliest TOI
//   rollback      refit at the rolled-back x, DCD at that time, append
// and the Daviet contact projection inside every ADMM iteration
// (solver/contact_constraints.cu:203-411).
//
// Device side: per object one implicit BVH over its triangles and one over
// its edges (Morton-sorted at every step start, refit every call), point and
// edge queries traverse them; the CCD kernel and the DCD kernel consume the
// candidate lists straight from device memory.  Host side: the DCD records
// are keyed on global vertex ids, sorted, de-duplicated against the contacts
// of the step and turned into Daviet contacts; the device colours them
// first-fit in contact order.  Both make the contact solve independent of
// the order in which the GPU found the pairs.
#include "api/collision.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <unordered_set>

#include "api/context.hpp"
#include "hip/collision_kernels.hpp"
#include "hip/kernels.hpp"

namespace silk::amd {

namespace {

struct ObjDesc {
  bool physical = false;
  bool self_on = false;
  int group = -1;  // -1: collision off (object_collider.cu:135,197)
  float pad = 0.0f;
  float friction = 0.0f;
  int v0 = 0, nv = 0;  // global vertex range
  int t0 = 0, nt = 0;  // global triangle range
  int e0 = 0, ne = 0;  // global edge range
  int tree_t = -1, tree_e = -1;
};

uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Primitives of nvert vertices each, ordered by the Morton code of their
// centroid (10 bits per axis over the centroid bounds).
std::vector<int> morton_order(const int* verts, int nprim, int nvert,
                              const std::vector<float>& pos) {
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  std::vector<std::array<float, 3>> cen(nprim);
  for (int i = 0; i < nprim; ++i) {
    for (int d = 0; d < 3; ++d) {
      float s = 0.0f;
      for (int k = 0; k < nvert; ++k) s += pos[3 * verts[nvert * i + k] + d];
      cen[i][d] = s / static_cast<float>(nvert);
      lo[d] = std::min(lo[d], cen[i][d]);
      hi[d] = std::max(hi[d], cen[i][d]);
    }
  }
  std::vector<std::pair<uint32_t, int>> code(nprim);
  for (int i = 0; i < nprim; ++i) {
    uint32_t c = 0;
    for (int d = 0; d < 3; ++d) {
      const float ext = hi[d] - lo[d];
      const float t = ext > 0.0f ? (cen[i][d] - lo[d]) / ext : 0.0f;
      const uint32_t q = static_cast<uint32_t>(std::clamp(t * 1023.0f, 0.0f, 1023.0f));
      c |= spread10(q) << (2 - d);
    }
    code[i] = {c, i};
  }
  std::sort(code.begin(), code.end());
  std::vector<int> order(nprim);
  for (int i = 0; i < nprim; ++i) order[i] = code[i].second;
  return order;
}

// Daviet contact in double, as assembled on the host (contact_constraints.cu:
// 23-150).
struct HostContact {
  int state[4];
  double coef[4], gamma[4], normal[3], affine[3], mu;
};

// Order-independent identity of a DCD record: type and global vertex ids,
// triangle corners sorted, each edge (min, max) and the smaller edge first.
struct KeyHash {
  size_t operator()(const std::array<int, 5>& k) const {
    uint64_t h = 1469598103934665603ull;
    for (int v : k) h = (h ^ static_cast<uint32_t>(v)) * 1099511628211ull;
    return static_cast<size_t>(h);
  }
};

// Pinned host staging buffer that grows; the caller makes sure no copy from
// it is in flight when it grows or is rewritten.
template <class T>
class PinnedVec {
 public:
  ~PinnedVec() {
    if (p_ != nullptr) (void)hipHostFree(p_);
  }
  T* reserve(size_t n) {
    if (n > cap_) {
      if (p_ != nullptr) (void)hipHostFree(p_);
      cap_ = std::max<size_t>(n, 2 * cap_);
      hip_check(hipHostMalloc(reinterpret_cast<void**>(&p_), cap_ * sizeof(T)),
                "hipHostMalloc");
    }
    return p_;
  }
  T* get() { return p_; }

 private:
  T* p_ = nullptr;
  size_t cap_ = 0;
};

std::array<int, 5> contact_key(const DcdRec& r) {
  std::array<int, 5> k{r.type, r.gv[0], r.gv[1], r.gv[2], r.gv[3]};
  if (r.type == 0) {
    std::sort(k.begin() + 2, k.end());
  } else {
    if (k[2] < k[1]) std::swap(k[1], k[2]);
    if (k[4] < k[3]) std::swap(k[3], k[4]);
    if (std::make_pair(k[3], k[4]) < std::make_pair(k[1], k[2])) {
      std::swap(k[1], k[3]);
      std::swap(k[2], k[4]);
    }
  }
  return k;
}

class HipCollision final : public CollisionWorld {
 public:
  HipCollision(Context& ctx, Scene& S, std::vector<ObjDesc> objs, int nvc,
               int nvt);
  ~HipCollision() override;

  void begin_step(const float* outer_x, float* ms, float* dhat,
                  silk_hip_step_stats* stats) override;
  float min_toi(const float* outer_x, const float* inner_x, float time_start,
                float interval, float ms, silk_hip_step_stats* stats) override;
  void rollback_contacts(const float* outer_x, float time, float dhat,
                         silk_hip_step_stats* stats) override;
  void contact_update(const float* x, const int* skip) override;
  ContactView contact_view() const override;
  int active_dof() const override { return 3 * n_active_; }
  void end_step(silk_hip_step_stats* stats) override;
  void debug_candidates(const float* pos0, const float* pos1,
                        std::vector<int>& pt, std::vector<int>& ee) override;
  int total_vertices() const override { return nvt_; }
  void debug_dcd(const float* pos0, const float* pos1, float time, float dhat,
                 std::vector<int>& recs, std::vector<float>& dist) override;
  const std::vector<int>& decision_log() const override { return log_; }
  const std::vector<int>& decision_info() const override { return info_; }
  int split_depth() const override { return split_depth_; }
  void set_split_depth(int depth) override;
  float time_broadphase(int reps) override;

  // Static tables (filled by make_collision_world).
  std::vector<int> h_tri, h_edge, h_tri_obj, h_edge_obj, h_point_obj;
  void finish_setup(const std::vector<float>& pos0);

 private:
  Context& ctx_;
  Scene& S_;
  hipStream_t st_;
  std::vector<ObjDesc> objs_;
  int nvc_, nvt_;
  float min_pad_ = 0.0f;
  float cur_dhat_ = 0.0f;

  DevBuf<float> p0_, p1_, invm_;
  DevBuf<int> tri_, edge_, point_, tri_obj_, edge_obj_, point_obj_;
  DevBuf<float> obj_friction_;
  std::vector<TreeDev> trees_;
  DevBuf<TreeDev> trees_dev_;
  // First refit workgroup of every tree (256 leaves each), then the total.
  DevBuf<int> refit_off_;
  int refit_blocks_ = 0;
  DevBuf<int> tree_prim_, tree_prim_id_;
  DevBuf<Box6> sbox_, node_;
  // Broadphase passes (object pairs, then self) in one launch.
  DevBuf<int> query_order_;  // per object: point order, then edge order
  std::vector<TraverseJob> jobs_h_;
  DevBuf<TraverseJob> jobs_;
  int job_items_ = 0;
  std::vector<int> job_height_;  // per job: its tree's height
  int split_depth_ = kTraverseSplitDepth;

  // Per-step Morton re-ordering of tree primitives and query orders
  // (SILK_HIP_BVH_REBUILD=0 keeps the set-up order).
  bool rebuild_order_ = true;
  int order_total_ = 0, order_nseg_ = 0;
  std::vector<size_t> point_off_, edge_off_;  // per object, into query_order_
  DevBuf<OrderSeg> order_segs_;
  DevBuf<int> order_off_, order_of_, ovals_in_, ovals_out_;
  DevBuf<unsigned long long> okeys_in_, okeys_out_;
  DevBuf<uint32_t> order_box_;
  DevBuf<int2> order_blk_;  // bounds workgroups: (segment, first item)
  int order_nblk_ = 0;
  DevBuf<char> order_temp_;
  size_t order_temp_bytes_ = 0;
  void setup_order_rebuild();
  void rebuild_order();

  // candidates and outputs
  DevBuf<int2> pt_pairs_, ee_pairs_;
  int pt_cap_ = 1 << 16, ee_cap_ = 1 << 16;
  DevBuf<DcdRec> recs_;
  int rec_cap_ = 1 << 12;
  int rec_guess_ = 256;          // records copied back with the counters
  PinnedVec<DcdRec> rec_pin_;
  DevBuf<int> counters_;  // pt, ee, dcd, pad (zeroed per broadphase)
  DevBuf<unsigned long long> toi_bits_;
  DevBuf<double> err6_;
  int* h_io_ = nullptr;  // pinned: counters[4] then toi bits (8 bytes)

  // contacts of the step
  int dcd_call_ = 0;                     // DCD calls so far in this step
  std::vector<int> log_;                 // appended contacts (decision_log)
  std::vector<int> info_;                // their colours and masks (decision_info)
  std::vector<HostContact> contacts_;
  std::unordered_set<std::array<int, 5>, KeyHash> keys_;
  std::vector<int> color_;               // per contact, from 1
  std::vector<uint8_t> solved_;          // per cloth vertex: a contact solves for it
  std::vector<int> touched_;             // vertices with contacts
  std::vector<int> cvert_;               // contact vertices, first use order
  std::vector<int> vloc_;                // per cloth vertex: index in cvert_
  // Independent groups of contacts (cloths coupled by no contact): contact
  // vertices numbered group by group (gloc_, the order of gvert_), the
  // contacts reordered by (group, colour) on the device, one workgroup per
  // group in the update when every group fits the register variant.
  std::vector<int> gloc_, gvert_, grp_h_, group_act_h_, group_n_h_;
  int ngroup_ = 1;
  bool grouped_ = false;
  bool groups_on_ = true;  // SILK_HIP_CONTACT_GROUPS=0 turns grouping off
  int groups_used_ = 0;    // largest grouped launch of the step
  DevBuf<int> grp_dev_, group_act_, group_off_, group_cnt_;
  DevBuf<ContactDev> c_grp_;  // (group, colour) order
  int max_color_ = 0;
  PinnedVec<ContactDev> hc_pin_;
  PinnedVec<int> col_pin_;  // old colours up, colouring results back
  PinnedVec<int> act_pin_;
  PinnedVec<float> w_pin_;
  std::vector<int> color_off_h_;
  int n_active_ = 0;
  DevBuf<ContactDev> c_dev_;  // colour order
  DevBuf<ContactDev> c_in_;   // contact order (colouring input)
  DevBuf<int> col_dev_, col_head_, col_meta_;
  DevBuf<unsigned> col_mask_;
  DevBuf<int> color_off_;
  DevBuf<float> corr_;
  int corr_cap_ = 0;
  DevBuf<float> w_, p_, eta_, dual_;
  DevBuf<int> act_;     // contact vertices (cvert_ on the device)
  DevBuf<float> pc_;    // 3 per contact vertex: scratch of the global variant
  int force_variant_ = -1;

  CollVerts verts() const;
  PrimTables tables() const;
  void set_cloth(const float* a, const float* b);
  void refit();
  void build_jobs(const std::vector<float>& pos0);
  void launch_broadphase();
  // reuse_pairs: test the candidate lists of the last broadphase (the CCD
  // of the segment this DCD's positions lie on) instead of a new one.
  std::vector<DcdRec> dcd(float time, float dhat, silk_hip_step_stats* stats,
                          bool reuse_pairs = false);
  void append(std::vector<DcdRec> recs);
  void color_and_upload(int n_old);
  void read_io();
};

HipCollision::HipCollision(Context& ctx, Scene& S, std::vector<ObjDesc> objs,
                           int nvc, int nvt)
    : ctx_(ctx), S_(S), st_(ctx.stream_), objs_(std::move(objs)), nvc_(nvc),
      nvt_(nvt) {
  hip_check(hipHostMalloc(reinterpret_cast<void**>(&h_io_), 32), "hipHostMalloc");
  std::memset(h_io_, 0, 32);
  min_pad_ = 1e30f;
  for (const auto& o : objs_) min_pad_ = std::min(min_pad_, o.pad);
}

HipCollision::~HipCollision() {
  if (h_io_ != nullptr) (void)hipHostFree(h_io_);
}

void HipCollision::finish_setup(const std::vector<float>& pos0) {
  const size_t n3 = 3 * static_cast<size_t>(nvt_);
  p0_.resize(n3);
  p1_.resize(n3);
  std::vector<float> hinv(nvt_, 0.0f);
  for (int i = 0; i < nvc_; ++i)
    hinv[i] = S_.h_pinned[i] ? 0.0f : static_cast<float>(1.0 / S_.h_mass[i]);
  invm_.upload(hinv, st_);
  std::vector<int> points(nvt_);
  for (int i = 0; i < nvt_; ++i) points[i] = i;
  std::vector<float> fr;
  for (const auto& o : objs_) fr.push_back(o.friction);
  tri_.upload(h_tri.empty() ? std::vector<int>{0, 0, 0} : h_tri, st_);
  edge_.upload(h_edge.empty() ? std::vector<int>{0, 0} : h_edge, st_);
  point_.upload(points, st_);
  tri_obj_.upload(h_tri_obj.empty() ? std::vector<int>{0} : h_tri_obj, st_);
  edge_obj_.upload(h_edge_obj.empty() ? std::vector<int>{0} : h_edge_obj, st_);
  point_obj_.upload(h_point_obj, st_);
  obj_friction_.upload(fr, st_);

  // Trees: primitives Morton-sorted by their centroid at pos0.
  std::vector<int> prim, prim_id;
  size_t nbox = 0, nnode = 0;
  struct Pending {
    int nprim, nvert, P;
    float pad;
    size_t prim_off, box_off, node_off;
  };
  std::vector<Pending> pend;
  auto add_tree = [&](const int* verts, int nprim, int nvert, int gid0,
                      float pad) {
    const std::vector<int> order = morton_order(verts, nprim, nvert, pos0);
    Pending p;
    p.nprim = nprim;
    p.nvert = nvert;
    int leaves = std::max(1, (nprim + kLeafPrims - 1) / kLeafPrims);
    p.P = 1;
    while (p.P < leaves) p.P <<= 1;
    p.pad = pad;
    p.prim_off = prim.size();
    p.box_off = nbox;
    p.node_off = nnode;
    for (const int i : order) {
      for (int k = 0; k < nvert; ++k) prim.push_back(verts[nvert * i + k]);
      prim_id.push_back(gid0 + i);
    }
    nbox += std::max(1, nprim);
    nnode += 2 * static_cast<size_t>(p.P);
    pend.push_back(p);
    return static_cast<int>(pend.size()) - 1;
  };
  for (auto& o : objs_) {
    o.tree_t = add_tree(h_tri.data() + 3 * static_cast<size_t>(o.t0), o.nt, 3,
                        o.t0, o.pad);
    o.tree_e = add_tree(h_edge.data() + 2 * static_cast<size_t>(o.e0), o.ne, 2,
                        o.e0, o.pad);
  }
  if (prim.empty()) prim.push_back(0);
  if (prim_id.empty()) prim_id.push_back(0);
  tree_prim_.upload(prim, st_);
  tree_prim_id_.upload(prim_id, st_);
  sbox_.resize(std::max<size_t>(1, nbox));
  node_.resize(std::max<size_t>(1, nnode));
  size_t id_off = 0;
  for (const auto& p : pend) {
    TreeDev t;
    t.nprim = p.nprim;
    t.nvert = p.nvert;
    t.P = p.P;
    t.pad = p.pad;
    t.prim = tree_prim_.get() + p.prim_off;
    t.prim_id = tree_prim_id_.get() + id_off;
    t.sbox = sbox_.get() + p.box_off;
    t.node = node_.get() + p.node_off;
    id_off += p.nprim;
    trees_.push_back(t);
  }
  trees_dev_.upload(trees_, st_);
  std::vector<int> roff(trees_.size() + 1, 0);
  for (size_t i = 0; i < trees_.size(); ++i)
    roff[i + 1] = roff[i] + (trees_[i].P + 255) / 256;
  refit_off_.upload(roff, st_);
  refit_blocks_ = roff.back();
  build_jobs(pos0);

  pt_pairs_.resize(pt_cap_);
  ee_pairs_.resize(ee_cap_);
  recs_.resize(rec_cap_);
  counters_.resize(4);
  toi_bits_.resize(1);
  err6_.resize(6);
  const size_t nc3 = 3 * static_cast<size_t>(std::max(1, nvc_));
  w_.resize(std::max(1, nvc_));
  for (DevBuf<float>* b : {&p_, &eta_, &dual_}) b->resize(nc3);
  vloc_.assign(std::max(1, nvc_), -1);
  if (const char* e = std::getenv("SILK_HIP_CONTACT_GROUPS")) groups_on_ = e[0] != '0';
  // SILK_HIP_CONTACT_VARIANT forces a contact-update variant (0..3) where it
  // fits, so that tests can run every one of them.
  if (const char* e = std::getenv("SILK_HIP_CONTACT_VARIANT"))
    force_variant_ = std::clamp(std::atoi(e), -1, 3);
  if (const char* e = std::getenv("SILK_HIP_BVH_REBUILD")) rebuild_order_ = std::atoi(e) != 0;
  setup_order_rebuild();
  // Host vectors above are temporaries: drain the uploads before they go.
  hip_check(hipStreamSynchronize(st_), "collision setup");
}

// Sort segments: every tree's primitives, then every object's point and
// edge query orders.
void HipCollision::setup_order_rebuild() {
  std::vector<OrderSeg> segs;
  std::vector<int> off{0};
  auto add = [&](OrderSeg g) {
    if (g.n <= 0) return;
    g.off = off.back();
    segs.push_back(g);
    off.push_back(off.back() + g.n);
  };
  for (const auto& o : objs_) {
    const TreeDev& tt = trees_[o.tree_t];
    OrderSeg g;
    g.n = tt.nprim;
    g.nvert = 3;
    g.src = tri_.get() + 3 * static_cast<size_t>(o.t0);
    g.out_prim = const_cast<int*>(tt.prim);
    g.out_id = const_cast<int*>(tt.prim_id);
    g.gid0 = o.t0;
    add(g);
    const TreeDev& te = trees_[o.tree_e];
    g.n = te.nprim;
    g.nvert = 2;
    g.src = edge_.get() + 2 * static_cast<size_t>(o.e0);
    g.out_prim = const_cast<int*>(te.prim);
    g.out_id = const_cast<int*>(te.prim_id);
    g.gid0 = o.e0;
    add(g);
  }
  for (size_t i = 0; i < objs_.size(); ++i) {
    const ObjDesc& o = objs_[i];
    OrderSeg g;
    g.n = o.nv;
    g.nvert = 1;
    g.src = point_.get() + o.v0;
    g.out_order = query_order_.get() + point_off_[i];
    add(g);
    g.n = o.ne;
    g.nvert = 2;
    g.src = edge_.get() + 2 * static_cast<size_t>(o.e0);
    g.out_order = query_order_.get() + edge_off_[i];
    add(g);
  }
  order_nseg_ = static_cast<int>(segs.size());
  order_total_ = off.back();
  if (order_total_ == 0) return;
  std::vector<int> of(order_total_);
  for (int k = 0; k < order_nseg_; ++k)
    for (int i = off[k]; i < off[k + 1]; ++i) of[i] = k;
  order_segs_.upload(segs, st_);
  order_off_.upload(off, st_);
  order_of_.upload(of, st_);
  order_box_.resize(6 * static_cast<size_t>(order_nseg_));
  std::vector<int2> blk;
  for (int k = 0; k < order_nseg_; ++k)
    for (int i = 0; i < segs[k].n; i += kOrderItemsPerBlock) blk.push_back(int2{k, i});
  order_nblk_ = static_cast<int>(blk.size());
  order_blk_.upload(blk, st_);
  okeys_in_.resize(order_total_);
  okeys_out_.resize(order_total_);
  ovals_in_.resize(order_total_);
  ovals_out_.resize(order_total_);
  order_temp_bytes_ = order_sort_temp_bytes(order_total_, order_nseg_);
  order_temp_.resize(std::max<size_t>(1, order_temp_bytes_));
  // The uploads above read host vectors that end with this function.
  hip_check(hipStreamSynchronize(st_), "order tables");
}

void HipCollision::rebuild_order() {
  if (order_total_ == 0) return;
  OrderScratch sc;
  sc.box = order_box_.get();
  sc.blk = order_blk_.get();
  sc.nblk = order_nblk_;
  sc.keys_in = okeys_in_.get();
  sc.keys_out = okeys_out_.get();
  sc.vals_in = ovals_in_.get();
  sc.vals_out = ovals_out_.get();
  sc.temp = order_temp_.get();
  sc.temp_bytes = order_temp_bytes_;
  hip_check(launch_order_rebuild(st_, p1_.get(), order_segs_.get(), order_nseg_,
                                 order_off_.get(), order_of_.get(), order_total_, sc),
            "BVH re-ordering");
}

CollVerts HipCollision::verts() const {
  CollVerts v;
  v.nvc = nvc_;
  v.nvt = nvt_;
  v.p0 = p0_.get();
  v.p1 = p1_.get();
  v.invm = invm_.get();
  return v;
}

PrimTables HipCollision::tables() const {
  PrimTables t;
  t.tri = tri_.get();
  t.edge = edge_.get();
  t.point = point_.get();
  t.tri_obj = tri_obj_.get();
  t.edge_obj = edge_obj_.get();
  t.point_obj = point_obj_.get();
  t.obj_friction = obj_friction_.get();
  return t;
}

void HipCollision::set_cloth(const float* a, const float* b) {
  const int64_t n = 3 * static_cast<int64_t>(nvc_);
  launch_copy_unless(st_, p0_.get(), a, n, nullptr);
  launch_copy_unless(st_, p1_.get(), b, n, nullptr);
  hip_check(hipGetLastError(), "collider positions");
}

void HipCollision::refit() {
  launch_refit(st_, verts(), trees_dev_.get(), static_cast<int>(trees_.size()),
               refit_off_.get(), refit_blocks_);
}

// Object pairs and self pairs of find_collision.cu:17-26,45-166: for each
// pair of objects in one group with at least one physical, the triangle tree
// of each against the other's points and the edge tree of the first against
// the second's edges, gated by the two object boxes; then each physical
// object with self-collision against itself.
void HipCollision::build_jobs(const std::vector<float>& pos0) {
  std::vector<int> order;
  std::vector<size_t>& point_off = point_off_;
  std::vector<size_t>& edge_off = edge_off_;
  point_off.assign(objs_.size(), 0);
  edge_off.assign(objs_.size(), 0);
  std::vector<int> ids;
  for (size_t i = 0; i < objs_.size(); ++i) {
    const ObjDesc& o = objs_[i];
    ids.resize(o.nv);
    for (int k = 0; k < o.nv; ++k) ids[k] = o.v0 + k;
    point_off[i] = order.size();
    const std::vector<int> po = morton_order(ids.data(), o.nv, 1, pos0);
    order.insert(order.end(), po.begin(), po.end());
    edge_off[i] = order.size();
    const std::vector<int> eo = morton_order(
        h_edge.data() + 2 * static_cast<size_t>(o.e0), o.ne, 2, pos0);
    order.insert(order.end(), eo.begin(), eo.end());
  }
  if (order.empty()) order.push_back(0);
  query_order_.upload(order, st_);
  auto points = [&](size_t i) {
    const ObjDesc& o = objs_[i];
    QuerySet q;
    q.n = o.nv;
    q.nvert = 1;
    q.prim = point_.get() + o.v0;
    q.order = query_order_.get() + point_off[i];
    q.pad = o.pad;
    q.gid0 = o.v0;
    return q;
  };
  auto edges = [&](size_t i) {
    const ObjDesc& o = objs_[i];
    QuerySet q;
    q.n = o.ne;
    q.nvert = 2;
    q.prim = edge_.get() + 2 * static_cast<size_t>(o.e0);
    q.order = query_order_.get() + edge_off[i];
    q.pad = o.pad;
    q.gid0 = o.e0;
    return q;
  };
  int items = 0;
  auto add = [&](const TreeDev& t, const QuerySet& q, int kind, int self,
                 const Box6* ga, const Box6* gb) {
    if (q.n == 0 || t.nprim == 0) return;
    TraverseJob j;
    j.tree = t;
    j.q = q;
    j.kind = kind;
    j.self = self;
    int height = 0;
    while ((1 << height) < t.P) ++height;
    j.depth = std::min(split_depth_, height);
    j.item0 = items;
    j.nitem = q.n << j.depth;
    j.gate_a = ga;
    j.gate_b = gb;
    items += (j.nitem + kTraverseWg - 1) / kTraverseWg * kTraverseWg;
    jobs_h_.push_back(j);
    job_height_.push_back(height);
  };
  auto root = [&](const ObjDesc& o) { return trees_[o.tree_t].node + 1; };
  const size_t n = objs_.size();
  for (size_t i = 0; i < n; ++i)
    for (size_t j = i + 1; j < n; ++j) {
      const ObjDesc& a = objs_[i];
      const ObjDesc& b = objs_[j];
      if (a.group == -1 || b.group == -1 || a.group != b.group) continue;
      if (!(a.physical || b.physical)) continue;
      add(trees_[a.tree_t], points(j), 0, 0, root(a), root(b));
      add(trees_[b.tree_t], points(i), 0, 0, root(a), root(b));
      add(trees_[a.tree_e], edges(j), 1, 0, root(a), root(b));
    }
  for (size_t i = 0; i < n; ++i) {
    const ObjDesc& o = objs_[i];
    if (!o.physical || !o.self_on) continue;
    add(trees_[o.tree_t], points(i), 0, 1, nullptr, nullptr);
    add(trees_[o.tree_e], edges(i), 1, 1, nullptr, nullptr);
  }
  job_items_ = items;
  if (!jobs_h_.empty()) jobs_.upload(jobs_h_, st_);
  // `order` is a temporary: drain the uploads before it goes.
  hip_check(hipStreamSynchronize(st_), "broadphase jobs");
}

void HipCollision::set_split_depth(int depth) {
  split_depth_ = std::max(0, depth);
  int items = 0;
  for (size_t k = 0; k < jobs_h_.size(); ++k) {
    TraverseJob& j = jobs_h_[k];
    j.depth = std::min(split_depth_, job_height_[k]);
    j.item0 = items;
    j.nitem = j.q.n << j.depth;
    items += (j.nitem + kTraverseWg - 1) / kTraverseWg * kTraverseWg;
  }
  job_items_ = items;
  if (!jobs_h_.empty()) jobs_.upload(jobs_h_, st_);
  hip_check(hipStreamSynchronize(st_), "broadphase jobs");
}

float HipCollision::time_broadphase(int reps) {
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  launch_broadphase();  // warm
  hip_check(hipEventRecord(e0, st_), "event");
  for (int k = 0; k < reps; ++k) launch_broadphase();
  hip_check(hipEventRecord(e1, st_), "event");
  hip_check(hipEventSynchronize(e1), "event sync");
  float ms = 0.0f;
  hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
  hip_check(hipEventDestroy(e0), "event");
  hip_check(hipEventDestroy(e1), "event");
  return ms / static_cast<float>(std::max(1, reps));
}

void HipCollision::launch_broadphase() {
  launch_fill_u32(st_, counters_.get(), 4, 0u);
  PairOut pt{pt_pairs_.get(), counters_.get(), pt_cap_};
  PairOut ee{ee_pairs_.get(), counters_.get() + 1, ee_cap_};
  launch_traverse_jobs(st_, verts(), jobs_.get(), static_cast<int>(jobs_h_.size()),
                       job_items_, pt, ee);
}

void HipCollision::read_io() {
  hip_check(hipMemcpyAsync(h_io_, counters_.get(), 4 * sizeof(int),
                           hipMemcpyDeviceToHost, st_),
            "counters");
  hip_check(hipMemcpyAsync(h_io_ + 4, toi_bits_.get(), sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, st_),
            "toi");
  hip_check(hipStreamSynchronize(st_), "collision sync");
}

static int grow(int need) { return std::max(1024, need + need / 2); }

std::vector<DcdRec> HipCollision::dcd(float time, float dhat,
                                      silk_hip_step_stats* stats, bool reuse_pairs) {
  for (;;) {
    if (reuse_pairs) {
      // Record count (and the unused slot after it) back to zero; the pair
      // counts stay.
      launch_fill_u32(st_, counters_.get() + 2, 2, 0u);
    } else {
      ctx_.timed_begin(kTimeBroad);
      launch_broadphase();
      ctx_.timed_end();
    }
    launch_dcd(st_, verts(), tables(), pt_pairs_.get(), counters_.get(), pt_cap_,
               ee_pairs_.get(), counters_.get() + 1, ee_cap_, time, dhat,
               recs_.get(), counters_.get() + 2, rec_cap_);
    // The records expected from the last call's count travel with the
    // counters, which saves a second round trip when they suffice.
    const int guess = std::min(rec_cap_, rec_guess_);
    DcdRec* hr = rec_pin_.reserve(static_cast<size_t>(std::max(1, guess)));
    hip_check(hipMemcpyAsync(hr, recs_.get(), guess * sizeof(DcdRec), hipMemcpyDeviceToHost,
                             st_),
              "dcd records");
    read_io();
    ctx_.timed_collect();
    const int npt = h_io_[0], nee = h_io_[1], nrec = h_io_[2];
    bool again = false;
    if (npt > pt_cap_) {
      pt_cap_ = grow(npt);
      pt_pairs_.resize(pt_cap_);
      again = true;
    }
    if (nee > ee_cap_) {
      ee_cap_ = grow(nee);
      ee_pairs_.resize(ee_cap_);
      again = true;
    }
    if (nrec > rec_cap_) {
      rec_cap_ = grow(nrec);
      recs_.resize(rec_cap_);
      again = true;
    }
    if (again) continue;
    if (stats) stats->dcd_candidates += npt + nee;
    std::vector<DcdRec> out(nrec);
    std::copy(hr, hr + std::min(nrec, guess), out.begin());
    if (nrec > guess) {
      DcdRec* rest = rec_pin_.reserve(static_cast<size_t>(nrec - guess));
      hip_check(hipMemcpyAsync(rest, recs_.get() + guess, (nrec - guess) * sizeof(DcdRec),
                               hipMemcpyDeviceToHost, st_),
                "dcd records");
      hip_check(hipStreamSynchronize(st_), "dcd sync");
      std::copy(rest, rest + (nrec - guess), out.begin() + guess);
    }
    rec_guess_ = std::max(64, nrec + nrec / 4);
    return out;
  }
}

// Contact assembly of contact_constraints.cu:23-150, in double.
void HipCollision::append(std::vector<DcdRec> recs) {
  const auto t0 = std::chrono::steady_clock::now();
  struct HostTime {  // records on every return path
    Context& c;
    std::chrono::steady_clock::time_point t0;
    ~HostTime() {
      if (!c.timing(kTimeHostContacts)) return;
      c.timing_ms_[kTimeHostContacts] +=
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
              .count();
      c.timing_launch_[kTimeHostContacts] += 1;
    }
  } host_time{ctx_, t0};
  std::sort(recs.begin(), recs.end(), [](const DcdRec& a, const DcdRec& b) {
    return contact_key(a) < contact_key(b);
  });
  const int n_old = static_cast<int>(contacts_.size());
  const int call = dcd_call_++;
  for (const DcdRec& r : recs) {
    if (!keys_.insert(contact_key(r)).second) continue;
    log_.insert(log_.end(), {call, r.type, r.gv[0], r.gv[1], r.gv[2], r.gv[3]});
    HostContact c;
    c.mu = r.friction;
    double ref[3] = {0.0, 0.0, 0.0};
    const double dhat = static_cast<double>(cur_dhat_);
    for (int d = 0; d < 3; ++d) {
      c.normal[d] = r.normal[d];
      c.affine[d] = -dhat * c.normal[d];
    }
    double scatter = 0.0;
    for (int k = 0; k < 4; ++k) {
      c.coef[k] = r.coef[k];
      for (int d = 0; d < 3; ++d) ref[d] += c.coef[k] * r.x[k][d];
      if (r.gv[k] >= nvc_) {
        c.state[k] = -1;
        for (int d = 0; d < 3; ++d) c.affine[d] += c.coef[k] * r.x[k][d];
        continue;
      }
      c.state[k] = r.gv[k];
      if (r.invm[k] > 0.0f)
        scatter += c.coef[k] * c.coef[k] / S_.h_contact_w[r.gv[k]];
    }
    for (int d = 0; d < 3; ++d) c.affine[d] += r.presc[d];
    const double rn =
        ref[0] * c.normal[0] + ref[1] * c.normal[1] + ref[2] * c.normal[2];
    for (int d = 0; d < 3; ++d) c.affine[d] -= ref[d] - rn * c.normal[d];
    for (int k = 0; k < 4; ++k) {
      if (c.state[k] >= 0 && r.invm[k] > 0.0f)
        c.gamma[k] = c.coef[k] / (scatter * S_.h_contact_w[c.state[k]]);
      else
        c.gamma[k] = 0.0;
    }
    contacts_.push_back(c);
  }
  if (static_cast<int>(contacts_.size()) == n_old) return;
  color_and_upload(n_old);
}

// Contact weights and contact vertices of the contacts appended since n_old,
// their colours from the device (k_color_contacts: first-fit in contact
// order over shared movable vertices, the order-independent equivalent of
// the sequential greedy colouring), and the contacts in colour order.
void HipCollision::color_and_upload(int n_old) {
  const int n = static_cast<int>(contacts_.size());
  if (solved_.size() < static_cast<size_t>(nvc_)) solved_.assign(nvc_, 0);
  float* hw = w_pin_.reserve(std::max(1, nvc_));
  if (n_old == 0 && touched_.empty())
    std::fill(hw, hw + std::max(1, nvc_), 0.0f);
  for (int c = n_old; c < n; ++c) {
    const HostContact& hc = contacts_[c];
    for (int k = 0; k < 4; ++k) {
      const int v = hc.state[k];
      if (hc.gamma[k] != 0.0 && !solved_[v]) {
        solved_[v] = 1;
        touched_.push_back(v);
        hw[v] = static_cast<float>(S_.h_contact_w[v]);
      }
      // Contact vertices, numbered in order of first use.
      if (v >= 0 && vloc_[v] < 0) {
        vloc_[v] = static_cast<int>(cvert_.size());
        cvert_.push_back(v);
      }
    }
  }
  n_active_ = static_cast<int>(touched_.size());
  const int nact = static_cast<int>(cvert_.size());
  // Groups: cloths joined by a contact (union-find over cloths), numbered
  // by their lowest cloth; contact vertices renumbered group by group.
  const int ncl = static_cast<int>(S_.vert_off.size()) - 1;
  auto cloth_of = [&](int v) {
    return static_cast<int>(std::upper_bound(S_.vert_off.begin(), S_.vert_off.end(), v) -
                            S_.vert_off.begin()) - 1;
  };
  std::vector<int> uf(std::max(1, ncl));
  for (int i = 0; i < ncl; ++i) uf[i] = i;
  auto find = [&](int a) {
    while (uf[a] != a) a = uf[a] = uf[uf[a]];
    return a;
  };
  for (int c = 0; c < n; ++c) {
    int first = -1;
    for (int k = 0; k < 4; ++k) {
      const int v = contacts_[c].state[k];
      if (v < 0) continue;
      const int r = find(cloth_of(v));
      if (first < 0) first = r;
      else if (r != first) uf[std::max(r, first)] = std::min(r, first);
    }
  }
  std::vector<int> gid(std::max(1, ncl), -1);
  ngroup_ = 0;
  std::vector<int> vgroup(nact);
  for (int l = 0; l < nact; ++l) {
    const int r = find(cloth_of(cvert_[l]));
    if (gid[r] < 0) gid[r] = 0;  // marks a cloth with contacts
  }
  for (int r = 0; r < ncl; ++r)
    if (gid[r] == 0 && find(r) == r) gid[r] = ngroup_++ + 1;
  for (int l = 0; l < nact; ++l) vgroup[l] = gid[find(cloth_of(cvert_[l]))] - 1;
  group_act_h_.assign(ngroup_ + 1, 0);
  for (int l = 0; l < nact; ++l) ++group_act_h_[vgroup[l] + 1];
  for (int g = 0; g < ngroup_; ++g) group_act_h_[g + 1] += group_act_h_[g];
  gvert_.assign(nact, 0);
  if (gloc_.size() < static_cast<size_t>(std::max(1, nvc_))) gloc_.assign(std::max(1, nvc_), -1);
  {
    std::vector<int> at(group_act_h_.begin(), group_act_h_.end() - 1);
    for (int l = 0; l < nact; ++l) {
      const int q = at[vgroup[l]]++;
      gvert_[q] = cvert_[l];
      gloc_[cvert_[l]] = q;
    }
  }
  grp_h_.assign(std::max(1, n), 0);
  group_n_h_.assign(std::max(1, ngroup_), 0);
  for (int c = 0; c < n; ++c) {
    int v = -1;
    for (int k = 0; k < 4 && v < 0; ++k) v = contacts_[c].state[k];
    grp_h_[c] = v >= 0 ? gid[find(cloth_of(v))] - 1 : 0;
    ++group_n_h_[grp_h_[c]];
  }
  int* act = act_pin_.reserve(std::max(1, nact));
  std::copy(gvert_.begin(), gvert_.end(), act);
  act_.reserve(std::max(1, nact));
  hip_check(hipMemcpyAsync(act_.get(), act, nact * sizeof(int), hipMemcpyHostToDevice,
                           st_),
            "contact vertices");
  pc_.reserve(3 * static_cast<size_t>(std::max(1, nact)));
  // Staging copy in contact order; the colouring kernel writes the colour
  // ordered one.  (No copy from the pinned buffers is in flight: the last
  // call ended with a synchronisation.)
  ContactDev* hc = hc_pin_.reserve(std::max(1, n));
  for (int i = 0; i < n; ++i) {
    const HostContact& c = contacts_[i];
    ContactDev& d = hc[i];
    for (int k = 0; k < 4; ++k) {
      d.state[k] = c.state[k] >= 0 ? gloc_[c.state[k]] : -1;
      // A static vertex's term is in affine; its coefficient is 0 here so
      // that the kernels add every term without a test.
      d.coef[k] = c.state[k] >= 0 ? static_cast<float>(c.coef[k]) : 0.0f;
      // gamma != 0 marks the slots the contact solves for (movable
      // vertices), also for the colouring.
      d.gamma[k] = static_cast<float>(c.gamma[k]);
    }
    for (int j = 0; j < 3; ++j) {
      d.normal[j] = static_cast<float>(c.normal[j]);
      d.affine[j] = static_cast<float>(c.affine[j]);
    }
    d.mu = static_cast<float>(c.mu);
    d.id = i;
  }
  c_in_.reserve(std::max(1, n));
  hip_check(hipMemcpyAsync(c_in_.get(), hc, n * sizeof(ContactDev),
                           hipMemcpyHostToDevice, st_),
            "contacts");
  const int n_new = n - n_old;
  constexpr int kBack = 3 + kMaxContactColors + 1;  // meta, then range ends
  int* hcol = col_pin_.reserve(static_cast<size_t>(n_old) + kBack + n_new);
  std::copy(color_.begin(), color_.begin() + n_old, hcol);
  col_dev_.reserve(std::max(1, n));
  if (n_old > 0)
    hip_check(hipMemcpyAsync(col_dev_.get(), hcol, n_old * sizeof(int),
                             hipMemcpyHostToDevice, st_),
              "colours");
  col_mask_.reserve(static_cast<size_t>(std::max(1, nact)) * kColorWords);
  col_head_.reserve(std::max(1, nact));
  c_dev_.reserve(std::max(1, n));
  color_off_.reserve(kMaxContactColors + 1);
  col_meta_.reserve(3);
  ColorArgs ca;
  ca.cin = c_in_.get();
  ca.n_old = n_old;
  ca.n = n;
  ca.nact = nact;
  ca.color = col_dev_.get();
  ca.mask = col_mask_.get();
  ca.head = col_head_.get();
  ca.cout = c_dev_.get();
  ca.off = color_off_.get();
  ca.meta = col_meta_.get();
  hip_check(launch_color_contacts(st_, ca), "contact colouring");
  // Reorder by (group, colour) right behind the colouring; whether the
  // update uses that order is decided once the colour count is known.
  const bool try_groups = groups_on_ && ngroup_ > 1;
  if (try_groups) {
    grp_dev_.reserve(std::max(1, n));
    hip_check(hipMemcpyAsync(grp_dev_.get(), grp_h_.data(), n * sizeof(int),
                             hipMemcpyHostToDevice, st_),
              "contact groups");
    group_act_.reserve(ngroup_ + 1);
    hip_check(hipMemcpyAsync(group_act_.get(), group_act_h_.data(),
                             (ngroup_ + 1) * sizeof(int), hipMemcpyHostToDevice, st_),
              "group vertices");
    const size_t nkey = static_cast<size_t>(ngroup_) * kMaxContactColors;
    group_cnt_.reserve(nkey);
    group_off_.reserve(nkey + 1);
    c_grp_.reserve(std::max(1, n));
    GroupArgs ga;
    ga.cin = c_dev_.get();
    ga.n = n;
    ga.ngroup = ngroup_;
    ga.group = grp_dev_.get();
    ga.color = col_dev_.get();
    ga.meta = col_meta_.get();
    ga.count = group_cnt_.get();
    ga.cout = c_grp_.get();
    ga.off = group_off_.get();
    hip_check(launch_group_contacts(st_, ga), "contact groups");
  }
  int* back = hcol + n_old;
  hip_check(hipMemcpyAsync(back, col_meta_.get(), 3 * sizeof(int), hipMemcpyDeviceToHost,
                           st_),
            "colouring result");
  hip_check(hipMemcpyAsync(back + 3, color_off_.get(), (kMaxContactColors + 1) * sizeof(int),
                           hipMemcpyDeviceToHost, st_),
            "colour ranges");
  hip_check(hipMemcpyAsync(back + kBack, col_dev_.get() + n_old, n_new * sizeof(int),
                           hipMemcpyDeviceToHost, st_),
            "new colours");
  // Staging buffers are not in flight after this: the weights below are
  // uploaded after the synchronisation.
  hip_check(hipStreamSynchronize(st_), "contact colouring");
  if (back[2] != 0)
    throw std::runtime_error("contact colouring needs more than " +
                             std::to_string(kMaxContactColors) + " colours");
  const int ncol = back[0];
  max_color_ = ncol;
  color_off_h_.assign(back + 3, back + 3 + ncol + 1);
  grouped_ = false;
  if (try_groups) {
    grouped_ = true;
    for (int g = 0; g < ngroup_ && grouped_; ++g)
      grouped_ = contact_group_fits(group_act_h_[g + 1] - group_act_h_[g], group_n_h_[g], ncol);
  }
  color_.resize(n);
  for (int c = n_old; c < n; ++c) {
    const int col = back[kBack + (c - n_old)];
    color_[c] = col;
    const HostContact& h = contacts_[c];
    int mask = 0;
    for (int k = 0; k < 4; ++k) mask |= h.gamma[k] != 0.0 ? 1 << k : 0;
    info_.push_back(col << 4 | mask);
  }
  hip_check(hipMemcpyAsync(w_.get(), hw, std::max(1, nvc_) * sizeof(float),
                           hipMemcpyHostToDevice, st_),
            "contact weights");
  // Corrections persist across appends; new contacts start at zero.
  if (n > corr_cap_) {
    const int cap = std::max(1024, 2 * n);
    DevBuf<float> nc(3 * static_cast<size_t>(cap));
    if (n_old > 0)
      hip_check(hipMemcpyAsync(nc.get(), corr_.get(), 3 * sizeof(float) * n_old,
                               hipMemcpyDeviceToDevice, st_),
                "corr grow");
    corr_ = std::move(nc);
    corr_cap_ = cap;
  }
  hip_check(hipMemsetAsync(corr_.get() + 3 * static_cast<size_t>(n_old), 0,
                           3 * sizeof(float) * (n - n_old), st_),
            "corr tail");
}

void HipCollision::begin_step(const float* outer_x, float* ms, float* dhat,
                              silk_hip_step_stats* stats) {
  // Obstacle segment prev -> curr of this frame.
  const size_t nvo = static_cast<size_t>(nvt_ - nvc_);
  if (nvo > 0) {
    std::vector<float> a, b;
    a.reserve(3 * nvo);
    b.reserve(3 * nvo);
    for (const auto& o : ctx_.obstacles_) {
      const std::vector<float>& cur = o.has_pos ? o.curr : o.mesh.V;
      const std::vector<float>& prv = o.has_pos ? o.prev : o.mesh.V;
      a.insert(a.end(), prv.begin(), prv.end());
      b.insert(b.end(), cur.begin(), cur.end());
    }
    hip_check(hipMemcpyAsync(p0_.get() + 3 * static_cast<size_t>(nvc_), a.data(),
                             a.size() * sizeof(float), hipMemcpyHostToDevice, st_),
              "obstacle t0");
    hip_check(hipMemcpyAsync(p1_.get() + 3 * static_cast<size_t>(nvc_), b.data(),
                             b.size() * sizeof(float), hipMemcpyHostToDevice, st_),
              "obstacle t1");
    hip_check(hipStreamSynchronize(st_), "obstacle upload");
  }
  // Minimum separation and DCD activation distance from the smallest
  // collider padding (main_loop.cu:186-200).
  *ms = ctx_.ccd_ms_scale() * min_pad_;
  *dhat = ctx_.dcd_scale() * min_pad_;
  cur_dhat_ = *dhat;
  // The contact set restarts every step.
  contacts_.clear();
  keys_.clear();
  log_.clear();
  info_.clear();
  dcd_call_ = 0;
  color_off_h_.clear();
  color_.clear();
  for (int v : touched_) {
    solved_[v] = 0;
    w_pin_.get()[v] = 0.0f;
  }
  touched_.clear();
  for (int v : cvert_) vloc_[v] = -1;
  cvert_.clear();
  max_color_ = 0;
  n_active_ = 0;
  const size_t nc3 = 3 * static_cast<size_t>(nvc_);
  launch_fill_u32(st_, w_.get(), nvc_, 0u);
  for (DevBuf<float>* b : {&p_, &eta_, &dual_})
    launch_fill_u32(st_, b->get(), static_cast<int64_t>(nc3), 0u);
  set_cloth(outer_x, outer_x);
  if (rebuild_order_) rebuild_order();
  refit();
  append(dcd(0.0f, *dhat, stats));
}

float HipCollision::min_toi(const float* outer_x, const float* inner_x,
                            float time_start, float interval, float ms,
                            silk_hip_step_stats* stats) {
  set_cloth(outer_x, inner_x);
  refit();
  launch_ccd_filter(st_, verts(), err6_.get());
  CcdParams cp;
  cp.ms = ms;
  cp.tolerance = static_cast<double>(ctx_.ccd_tolerance()) / interval;
  cp.max_iter = ctx_.ccd_max_iter();
  cp.time_start = time_start;
  cp.interval = interval;
  cp.err_vf = err6_.get();
  cp.err_ee = err6_.get() + 3;
  cp.zero_mode = ctx_.ccd_zero_mode();
  // Counter 3 (cleared with the pair counters by every broadphase launch).
  cp.zero_count = counters_.get() + 3;
  for (;;) {
    const double one = 1.0;
    unsigned long long bits;
    std::memcpy(&bits, &one, sizeof(bits));
    // toi bits start at 1.0.
    launch_fill_u64(st_, toi_bits_.get(), 1, bits);
    ctx_.timed_begin(kTimeBroad);
    launch_broadphase();
    ctx_.timed_end();
    ctx_.timed_begin(kTimeCcd);
    launch_ccd(st_, verts(), tables(), pt_pairs_.get(), counters_.get(), pt_cap_,
               ee_pairs_.get(), counters_.get() + 1, ee_cap_, cp,
               toi_bits_.get());
    ctx_.timed_end();
    read_io();
    ctx_.timed_collect();
    const int npt = h_io_[0], nee = h_io_[1];
    bool again = false;
    if (npt > pt_cap_) {
      pt_cap_ = grow(npt);
      pt_pairs_.resize(pt_cap_);
      again = true;
    }
    if (nee > ee_cap_) {
      ee_cap_ = grow(nee);
      ee_pairs_.resize(ee_cap_);
      again = true;
    }
    if (again) continue;
    if (stats) {
      stats->ccd_candidates += npt + nee;
      stats->ccd_zero_toi_pairs += h_io_[3];
    }
    double toi;
    std::memcpy(&toi, h_io_ + 4, sizeof(toi));
    return static_cast<float>(static_cast<double>(interval) * toi);
  }
}

// The DCD after a rollback tests the candidates of the CCD that found the
// rollback's TOI instead of a broadphase of its own (find_collision.cu runs
// one): the rollback positions lie on that CCD's segment (and obstacles
// within their frame motion), so every pair whose padded boxes overlap
// there is in the CCD's list of overlapping swept boxes, and the DCD's
// distance test selects the same records from the longer list.
void HipCollision::rollback_contacts(const float* outer_x, float time,
                                     float dhat, silk_hip_step_stats* stats) {
  set_cloth(outer_x, outer_x);
  append(dcd(time, dhat, stats, /*reuse_pairs=*/true));
}

void HipCollision::contact_update(const float* x, const int* skip) {
  if (contacts_.empty()) return;
  ContactSet cs;
  cs.n = static_cast<int>(contacts_.size());
  cs.ncolor = static_cast<int>(color_off_h_.size()) - 1;
  cs.nact = static_cast<int>(cvert_.size());
  cs.act = act_.get();
  cs.c = c_dev_.get();
  cs.color_off = color_off_.get();
  cs.corr = corr_.get();
  cs.w = w_.get();
  cs.p = p_.get();
  cs.eta = eta_.get();
  cs.dual = dual_.get();
  cs.pc = pc_.get();
  cs.force_variant = force_variant_;
  if (grouped_ && force_variant_ < 0) {
    cs.c = c_grp_.get();
    cs.ngroup = ngroup_;
    cs.group_off = group_off_.get();
    cs.group_act = group_act_.get();
    cs.group_n_max = *std::max_element(group_n_h_.begin(), group_n_h_.end());
    int am = 0;
    for (int g = 0; g < ngroup_; ++g) am = std::max(am, group_act_h_[g + 1] - group_act_h_[g]);
    cs.group_nact_max = am;
    groups_used_ = std::max(groups_used_, ngroup_);
  }
  cs.sweeps = ctx_.contact_sweeps();
  cs.skip = skip;
  launch_contact_update(st_, cs, x);
}

ContactView HipCollision::contact_view() const {
  ContactView v;
  if (contacts_.empty()) return v;
  v.w = w_.get();
  v.p = p_.get();
  v.eta = eta_.get();
  v.dual = dual_.get();
  return v;
}

void HipCollision::end_step(silk_hip_step_stats* stats) {
  const int used = groups_used_;
  groups_used_ = 0;
  if (!stats) return;
  stats->contact_groups = used;
  stats->contacts = static_cast<int>(contacts_.size());
  stats->contact_colors =
      color_off_h_.empty() ? 0 : static_cast<int>(color_off_h_.size()) - 1;
}

void HipCollision::debug_dcd(const float* pos0, const float* pos1, float time,
                             float dhat, std::vector<int>& recs,
                             std::vector<float>& dist) {
  const size_t bytes = 3 * static_cast<size_t>(nvt_) * sizeof(float);
  hip_check(hipMemcpyAsync(p0_.get(), pos0, bytes, hipMemcpyHostToDevice, st_),
            "debug t0");
  hip_check(hipMemcpyAsync(p1_.get(), pos1, bytes, hipMemcpyHostToDevice, st_),
            "debug t1");
  refit();
  const std::vector<DcdRec> out = dcd(time, dhat, nullptr);
  recs.clear();
  dist.clear();
  for (const DcdRec& r : out) {
    recs.insert(recs.end(), {r.type, r.gv[0], r.gv[1], r.gv[2], r.gv[3], 0});
    // Distance along the normal of the closest points: sum_k coef_k x_k
    // projected on the normal (the coefficients carry the signs).
    double d = 0.0;
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c)
        d += static_cast<double>(r.coef[k]) * r.x[k][c] * r.normal[c];
    dist.push_back(static_cast<float>(d));
  }
}

void HipCollision::debug_candidates(const float* pos0, const float* pos1,
                                    std::vector<int>& pt, std::vector<int>& ee) {
  const size_t bytes = 3 * static_cast<size_t>(nvt_) * sizeof(float);
  hip_check(hipMemcpyAsync(p0_.get(), pos0, bytes, hipMemcpyHostToDevice, st_),
            "debug t0");
  hip_check(hipMemcpyAsync(p1_.get(), pos1, bytes, hipMemcpyHostToDevice, st_),
            "debug t1");
  if (rebuild_order_) rebuild_order();
  refit();
  int npt = 0, nee = 0;
  for (;;) {
    launch_broadphase();
    read_io();
    npt = h_io_[0];
    nee = h_io_[1];
    bool again = false;
    if (npt > pt_cap_) {
      pt_cap_ = grow(npt);
      pt_pairs_.resize(pt_cap_);
      again = true;
    }
    if (nee > ee_cap_) {
      ee_cap_ = grow(nee);
      ee_pairs_.resize(ee_cap_);
      again = true;
    }
    if (!again) break;
  }
  std::vector<int2> hp(npt), he(nee);
  if (npt > 0)
    hip_check(hipMemcpyAsync(hp.data(), pt_pairs_.get(), npt * sizeof(int2),
                             hipMemcpyDeviceToHost, st_),
              "debug pt");
  if (nee > 0)
    hip_check(hipMemcpyAsync(he.data(), ee_pairs_.get(), nee * sizeof(int2),
                             hipMemcpyDeviceToHost, st_),
              "debug ee");
  hip_check(hipStreamSynchronize(st_), "debug sync");
  pt.clear();
  ee.clear();
  for (const int2& q : hp) {
    pt.push_back(q.y);
    for (int k = 0; k < 3; ++k) pt.push_back(h_tri[3 * static_cast<size_t>(q.x) + k]);
  }
  for (const int2& q : he) {
    int a0 = h_edge[2 * static_cast<size_t>(q.x)], a1 = h_edge[2 * static_cast<size_t>(q.x) + 1];
    int b0 = h_edge[2 * static_cast<size_t>(q.y)], b1 = h_edge[2 * static_cast<size_t>(q.y) + 1];
    const auto ka = std::make_pair(std::min(a0, a1), std::max(a0, a1));
    const auto kb = std::make_pair(std::min(b0, b1), std::max(b0, b1));
    if (kb < ka) {
      std::swap(a0, b0);
      std::swap(a1, b1);
    }
    ee.insert(ee.end(), {a0, a1, b0, b1});
  }
}

}  // namespace

std::unique_ptr<CollisionWorld> make_collision_world(Context& ctx, Scene& S) {
  // Objects: cloths (solver order, scene offsets), then obstacles.
  std::vector<ObjDesc> objs;
  std::vector<int> tri, edge, tri_obj, edge_obj, point_obj;
  int v0 = 0;
  for (size_t ci = 0; ci < ctx.cloths_.size(); ++ci) {
    const ClothEntity& c = ctx.cloths_[ci];
    ObjDesc o;
    o.physical = true;
    o.self_on = c.coll.is_self_collision_on != 0;
    o.group = c.coll.is_collision_on ? c.coll.group : -1;
    o.pad = 0.05f * c.mesh.avg_edge_length;
    o.friction = c.coll.friction;
    o.v0 = S.vert_off[ci];
    o.nv = c.mesh.nv();
    o.t0 = static_cast<int>(tri.size() / 3);
    o.nt = c.pmesh.nf();
    o.e0 = static_cast<int>(edge.size() / 2);
    o.ne = c.pmesh.ne();
    for (int f : c.pmesh.F) tri.push_back(o.v0 + f);
    for (int e : c.pmesh.E) edge.push_back(o.v0 + e);
    v0 = o.v0 + o.nv;
    objs.push_back(o);
  }
  const int nvc = v0;
  for (const auto& ob : ctx.obstacles_) {
    ObjDesc o;
    o.physical = false;
    o.self_on = ob.coll.is_self_collision_on != 0;
    o.group = ob.coll.is_collision_on ? ob.coll.group : -1;
    o.pad = 0.05f * ob.mesh.avg_edge_length;
    o.friction = ob.coll.friction;
    o.v0 = v0;
    o.nv = ob.mesh.nv();
    o.t0 = static_cast<int>(tri.size() / 3);
    o.nt = ob.mesh.nf();
    o.e0 = static_cast<int>(edge.size() / 2);
    o.ne = ob.mesh.ne();
    for (int f : ob.mesh.F) tri.push_back(o.v0 + f);
    for (int e : ob.mesh.E) edge.push_back(o.v0 + e);
    v0 += o.nv;
    objs.push_back(o);
  }
  const int nvt = v0;
  // Anything to collide at all?
  bool any = false;
  for (size_t i = 0; i < objs.size() && !any; ++i) {
    if (objs[i].physical && objs[i].self_on) any = true;
    for (size_t j = i + 1; j < objs.size() && !any; ++j)
      any = objs[i].group != -1 && objs[i].group == objs[j].group &&
            (objs[i].physical || objs[j].physical);
  }
  if (!any) return nullptr;
  for (size_t k = 0; k < objs.size(); ++k) {
    for (int t = 0; t < objs[k].nt; ++t) tri_obj.push_back(static_cast<int>(k));
    for (int e = 0; e < objs[k].ne; ++e) edge_obj.push_back(static_cast<int>(k));
    for (int v = 0; v < objs[k].nv; ++v) point_obj.push_back(static_cast<int>(k));
  }
  // Positions used to order the primitives along a Morton curve.
  std::vector<float> pos(3 * static_cast<size_t>(nvt));
  hip_check(hipMemcpyAsync(pos.data(), S.x.get(), 3 * sizeof(float) * nvc,
                           hipMemcpyDeviceToHost, ctx.stream_),
            "collider positions");
  hip_check(hipStreamSynchronize(ctx.stream_), "collider positions sync");
  {
    size_t off = 3 * static_cast<size_t>(nvc);
    for (const auto& ob : ctx.obstacles_) {
      const std::vector<float>& cur = ob.has_pos ? ob.curr : ob.mesh.V;
      std::copy(cur.begin(), cur.end(), pos.begin() + off);
      off += cur.size();
    }
  }
  auto hc = std::make_unique<HipCollision>(ctx, S, std::move(objs), nvc, nvt);
  hc->h_tri = std::move(tri);
  hc->h_edge = std::move(edge);
  hc->h_tri_obj = std::move(tri_obj);
  hc->h_edge_obj = std::move(edge_obj);
  hc->h_point_obj = std::move(point_obj);
  hc->finish_setup(pos);
  return hc;
}

}  // namespace silk::amd
