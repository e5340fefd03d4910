// Collision kernels: collider boxes, implicit BVH refit and traversal, DCD,
// FP64 CCD and the Daviet contact Gauss-Seidel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace silk::amd {

struct Box6 {
  float mn[3], mx[3];
};

/// Global collider vertex numbering: cloth vertices [0, nvc) in solver order,
/// then obstacle vertices.  p0/p1 hold the segment endpoints.
struct CollVerts {
  int nvc = 0, nvt = 0;
  const float* p0 = nullptr;
  const float* p1 = nullptr;
  const float* invm = nullptr;  // 0 for pinned / obstacle
};

/// One implicit BVH over a primitive set (triangles or edges of one object):
/// primitives sorted once by Morton code, leaves of kLeafPrims consecutive
/// sorted primitives, complete binary heap (root 1, leaves [P, 2P)).
constexpr int kLeafPrims = 4;
struct TreeDev {
  int nprim = 0;
  int nvert = 0;           // vertices per primitive (2 or 3)
  int P = 1;               // padded leaf count (power of two)
  float pad = 0.0f;
  const int* prim = nullptr;   // nvert * nprim global vertex ids (sorted order)
  const int* prim_id = nullptr;// original primitive id per sorted slot
  Box6* sbox = nullptr;        // nprim, sorted order
  Box6* node = nullptr;        // 2P
};

/// Refit all trees: 256 leaves per workgroup (leaf_block_off: ntrees + 1
/// prefix offsets of ceil(P / 256) workgroups per tree, nblocks in all),
/// then the levels above them, one workgroup per tree.
void launch_refit(hipStream_t s, const CollVerts& v, const TreeDev* trees_dev,
                  int ntrees, const int* leaf_block_off, int nblocks);

/// One sort segment of the per-step re-ordering: a tree's primitives (writes
/// out_prim and out_id) or an object's broadphase queries (writes out_order).
struct OrderSeg {
  int n = 0;                  // items
  int nvert = 1;              // vertices per item (1 point, 2 edge, 3 triangle)
  const int* src = nullptr;   // nvert * n global vertex ids, original order
  int* out_prim = nullptr;    // tree: nvert * n vertex ids in the new order
  int* out_id = nullptr;      // tree: n global primitive ids in the new order
  int* out_order = nullptr;   // queries: n local indices in the new order
  int gid0 = 0;               // global primitive id of item 0
  int off = 0;                // first key of the segment
};

struct OrderScratch {
  uint32_t* box = nullptr;     // 6 per segment: centroid bounds, order-preserving
                               //   integer encodings of the floats
  const int2* blk = nullptr;   // per bounds workgroup: (segment, first item)
  int nblk = 0;
  unsigned long long* keys_in = nullptr;  // total each: segment << 32 | code
  unsigned long long* keys_out = nullptr;
  int* vals_in = nullptr;
  int* vals_out = nullptr;
  void* temp = nullptr;        // radix sort workspace
  size_t temp_bytes = 0;
};

/// Items of one segment per workgroup of the centroid-bounds pass.
constexpr int kOrderItemsPerBlock = 2048;

/// Radix-sort workspace bytes for `total` keys of `nseg` segments.
size_t order_sort_temp_bytes(int total, int nseg);

/// Morton codes of every segment's items at positions p (3 per collider
/// vertex), one stable radix sort of (segment, code) keys, then the new
/// orders written (oibvh.cuh:521-560,603-633).  seg_off: nseg + 1 key
/// offsets; seg_of: the segment of every key.
hipError_t launch_order_rebuild(hipStream_t s, const float* p, const OrderSeg* segs, int nseg,
                                const int* seg_off, const int* seg_of, int total,
                                OrderScratch sc);

struct QuerySet {
  int n = 0;
  int nvert = 1;                 // 1 point, 2 edge
  const int* prim = nullptr;     // nvert * n global vertex ids
  const int* order = nullptr;    // n: visiting order of the queries (Morton)
  float pad = 0.0f;
  int gid0 = 0;                  // global primitive id of query 0
};

/// Pair output: (tree primitive global id, query primitive global id) plus a
/// running counter; entries beyond cap are counted but not written.
struct PairOut {
  int2* pairs = nullptr;
  int* count = nullptr;
  int cap = 0;
};

/// One tree x query-set pass of a broadphase.
/// kind 0: triangle tree x points (PT), 1: edge tree x edges (EE).
/// self: same object (adjacency filter; EE keeps tree id < query id only).
/// gate_a / gate_b: object boxes (triangle tree roots); when both are given
/// and do not overlap the pass reports nothing (find_collision.cu:45-92).
/// Each query is walked by 2^depth lanes, one per subtree at that depth;
/// the job's work items start at item0 (a multiple of kTraverseWg).
struct TraverseJob {
  TreeDev tree;
  QuerySet q;
  int kind = 0;
  int self = 0;
  int depth = 0;
  int item0 = 0;
  int nitem = 0;   // q.n << depth
  const Box6* gate_a = nullptr;
  const Box6* gate_b = nullptr;
};

/// Subtree split depth of a traversal job (clamped to the tree height).
constexpr int kTraverseSplitDepth = 5;
/// Threads per traversal workgroup; job item ranges are whole workgroups.
constexpr int kTraverseWg = 256;

/// All passes of a broadphase in one launch (nitems: end of the last job's
/// padded item range).  Pairs go to pt (kind 0) or ee (kind 1).
void launch_traverse_jobs(hipStream_t s, const CollVerts& v, const TraverseJob* jobs,
                          int njobs, int nitems, PairOut pt, PairOut ee);

/// DCD record of one active contact (dcd.cu:102-261), global vertex ids.
struct DcdRec {
  int type;     // 0 PT, 1 EE
  int gv[4];
  float coef[4];
  float invm[4];
  float normal[3];
  float x[4][3];
  float presc[3];
  float friction;   // sqrt(mu_a mu_b), rounded once from double
};

struct PrimTables {
  const int* tri = nullptr;    // 3 per triangle (global vertex ids)
  const int* edge = nullptr;   // 2 per edge
  const int* point = nullptr;  // 1 per point
  const int* tri_obj = nullptr;
  const int* edge_obj = nullptr;
  const int* point_obj = nullptr;
  const float* obj_friction = nullptr;
};

/// DCD over PT and EE candidate lists at fixed normalized time.
void launch_dcd(hipStream_t s, const CollVerts& v, const PrimTables& pt,
                const int2* pt_pairs, const int* pt_count, int pt_cap,
                const int2* ee_pairs, const int* ee_count, int ee_cap,
                float time, float dhat, DcdRec* out, int* out_count,
                int out_cap);

struct CcdParams {
  double ms = 0.0;
  double tolerance = 1e-6;   // already divided by the segment length
  int max_iter = 4096;
  double time_start = 0.0;
  double interval = 1.0;
  const double* err_vf = nullptr;  // device, 3 doubles
  const double* err_ee = nullptr;
  int zero_mode = 0;           // SILK_HIP_CCD_ZERO_SKIP / _CONTACT
  int* zero_count = nullptr;   // device: candidates left with a root at t = 0
};

/// Per-axis max |coordinate| over all collider vertices at t0 and t1, then the
/// VF / EE numerical filters (interval_root_finder.cpp:512-522) into err[6].
void launch_ccd_filter(hipStream_t s, const CollVerts& v, double* err6);

/// Pending-box capacity of one CCD query (overflow returns the current box's
/// t.lower: an earlier, conservative TOI).  One box per lane of a wavefront.
constexpr int kCcdStackCap = 64;

/// Workgroups (of four wavefronts) of a CCD launch when the occupancy query
/// fails; otherwise the launch holds as many as can be resident.
constexpr int kCcdBlocks = 1024;

/// Earliest TOI over the candidates (local units, written as the bit pattern
/// of a positive double into *toi_bits via atomicMin; initialise to 1.0).
void launch_ccd(hipStream_t s, const CollVerts& v, const PrimTables& pt,
                const int2* pt_pairs, const int* pt_count, int pt_cap,
                const int2* ee_pairs, const int* ee_count, int ee_cap,
                const CcdParams& cp, unsigned long long* toi_bits);

/// Per-query CCD on explicit FP64 end positions (12 doubles each for p0 and
/// p1, err 3 doubles per query) with the kernel of launch_ccd: hit[i] = 1 and
/// toi[i] as the step's CCD computes them under zero_mode (a root at t = 0
/// at the last separation searched is no hit).
void launch_ccd_each(hipStream_t s, int n, const int* is_vf, const double* p0,
                     const double* p1, const double* err, double ms,
                     double tolerance, int max_iter, int* hit, double* toi,
                     int zero_mode = 0);

/// Earliest TOI over explicit queries with the kernel of launch_ccd
/// (queries that cannot lower the running minimum are cut short); *toi_bits
/// starts at 1.0.
void launch_ccd_min(hipStream_t s, int n, const int* is_vf, const double* p0,
                    const double* p1, const double* err, double ms,
                    double tolerance, int max_iter, unsigned long long* toi_bits,
                    int zero_mode = 0);

// ---------------------------------------------------------------- contacts

/// The Coulomb projection of the contact kernels (contact_constraints.cu:
/// 391-411) on n explicit cases: u, nrm and out hold 3 floats per case.
void launch_coulomb_each(hipStream_t s, int n, const float* u, const float* nrm,
                         const float* mu, float* out);

/// Daviet contact (contact_constraints.cuh:19-42).
struct ContactDev {
  int state[4];     // contact vertex (index into ContactSet::act) or -1
  float coef[4];    // 0 where state is -1 (that term is in affine)
  float gamma[4];
  float normal[3];
  float affine[3];
  float mu;
  int id;           // contact id (index of its correction)
};

struct ContactSet {
  int n = 0;
  int ncolor = 0;
  // Contact vertices: the cloth vertices some contact references (movable
  // or not), in order of first use; the update works on these only.
  int nact = 0;
  const int* act = nullptr;         // nact cloth vertex ids (solver order)
  const ContactDev* c = nullptr;    // stored colour by colour
  const int* color_off = nullptr;   // ncolor + 1 offsets into c
  float* corr = nullptr;            // 3 per contact
  const float* w = nullptr;         // per cloth vertex: active weight (0 = inactive)
  float* p = nullptr;               // 3 per cloth vertex (read where w > 0)
  float* eta = nullptr;
  float* dual = nullptr;            // w (p - p_previous)
  float* pc = nullptr;              // 3 per contact vertex: scratch of the
                                    //   variant that keeps p out of LDS
  int sweeps = 10;
  const int* skip = nullptr;        // no-op when *skip != 0
  int force_variant = -1;           // >= 0: use that variant (tests)
  // Independent groups (no contact of one group shares a contact vertex
  // with another): with ngroup > 1, c is stored group by group and colour by
  // colour inside a group, the contact vertices group by group, and one
  // workgroup updates each group.
  int ngroup = 1;
  const int* group_off = nullptr;   // ngroup * ncolor + 1: start of each
                                    //   (group, colour) range, then the end
  const int* group_act = nullptr;   // ngroup + 1: contact-vertex range per group
  int group_n_max = 0;              // largest group: contacts
  int group_nact_max = 0;           //   and contact vertices
};

/// Device contact colouring (contact_coloring.cu:69-373): first-fit in
/// contact order over shared movable vertices, colours 1..kMaxContactColors.
constexpr int kColorWords = 8;
constexpr int kMaxContactColors = 32 * kColorWords - 1;
struct ColorArgs {
  const ContactDev* cin = nullptr;  // n contacts in contact order
  int n_old = 0;                    // contacts [0, n_old) keep their colours
  int n = 0;
  int nact = 0;                     // contact vertices (ContactDev::state range)
  int* color = nullptr;             // n: [0, n_old) read, [n_old, n) written
  unsigned* mask = nullptr;         // scratch: nact * kColorWords
  int* head = nullptr;              // scratch: nact
  ContactDev* cout = nullptr;       // n, colour by colour, id = contact index
  int* off = nullptr;               // kMaxContactColors + 1 range ends
  int* meta = nullptr;              // colours, rounds, overflow flag
};
hipError_t launch_color_contacts(hipStream_t s, const ColorArgs& a);

/// Contacts of independent groups: the colour-ordered contacts (cin, ids
/// set) reordered group by group, colour by colour inside a group.
struct GroupArgs {
  const ContactDev* cin = nullptr;  // n, colour by colour
  int n = 0;
  int ngroup = 0;
  const int* group = nullptr;       // per contact id: its group
  const int* color = nullptr;       // per contact id: its colour (1..)
  const int* meta = nullptr;        // colouring result (meta[0]: colours)
  int* count = nullptr;             // scratch: ngroup * kMaxContactColors
  ContactDev* cout = nullptr;       // n
  int* off = nullptr;               // ngroup * colours + 1 range starts
};
hipError_t launch_group_contacts(hipStream_t s, const GroupArgs& a);

/// One ADMM contact update (contact_constraints.cu:246-264): single
/// workgroup, colours separated by workgroup barriers.  p and the
/// corrections live in LDS for the whole update when they fit.
void launch_contact_update(hipStream_t s, const ContactSet& cs,
                           const float* x);
/// A group of nact contact vertices and n contacts fits the grouped update.
bool contact_group_fits(int nact, int n, int ncolor);

/// Variants of launch_contact_update, in the order it prefers them.
enum : int {
  kContactRegNm = 0,   // contacts in registers, normal and mu in LDS
  kContactReg = 1,     // contacts in registers
  kContactLds = 2,     // p and corrections in LDS
  kContactGlobal = 3,  // p and corrections in global memory
};

/// The variant launch_contact_update takes for nact contact vertices, n
/// contacts in ncolor colours.
int contact_update_variant(int nact, int n, int ncolor);

}  // namespace silk::amd
