// extern "C" entry points of include/silk_hip.h.  Every function converts
// exceptions into status codes; nothing unwinds across the ABI.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "api/context.hpp"
#include "cpu/cpu_world.hpp"
#include "hip/kernels.hpp"
#include "hip/collision_kernels.hpp"
#include "silk_hip.h"

using silk::amd::ClothEntity;
using silk::amd::CollisionWorld;
using silk::amd::Context;
using silk::amd::ObstacleEntity;
using silk::amd::TriMesh;

struct silk_hip_ctx {
  Context impl;
  explicit silk_hip_ctx(int d) : impl(d) {}
};

namespace {

template <typename F>
int guarded(silk_hip_ctx* ctx, F&& f) {
  if (ctx == nullptr) return SILK_HIP_ERR_NOT_INITIALIZED;
  ctx->impl.last_error_.clear();
  try {
    return f(ctx->impl);
  } catch (const std::bad_alloc&) {
    ctx->impl.last_error_ = "out of memory";
  } catch (const std::exception& e) {
    ctx->impl.last_error_ = e.what();
    if (std::string(e.what()) == "no HIP device") return SILK_HIP_ERR_NO_DEVICE;
  } catch (...) {
    ctx->impl.last_error_ = "unknown exception";
  }
  return SILK_HIP_ERR_UNKNOWN;
}

bool valid_pins(const int32_t* pins, int64_t npin, int nv) {
  if (npin < 0 || npin > nv) return false;
  if (npin > 0 && pins == nullptr) return false;
  for (int64_t i = 0; i < npin; ++i)
    if (pins[i] < 0 || pins[i] >= nv) return false;
  return true;
}

}  // namespace

extern "C" {

int silk_hip_device_count(int32_t* count) {
  if (count == nullptr) return SILK_HIP_ERR_INVALID_CONFIG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  *count = n;
  return SILK_HIP_OK;
}

int silk_hip_create(int32_t device, silk_hip_ctx** out) {
  if (out == nullptr) return SILK_HIP_ERR_INVALID_CONFIG;
  *out = nullptr;
  try {
    *out = new silk_hip_ctx(device);
  } catch (...) {
    return SILK_HIP_ERR_UNKNOWN;
  }
  return SILK_HIP_OK;
}

int silk_hip_destroy(silk_hip_ctx* ctx) {
  try {
    delete ctx;
  } catch (...) {
    return SILK_HIP_ERR_UNKNOWN;
  }
  return SILK_HIP_OK;
}

const char* silk_hip_last_error(const silk_hip_ctx* ctx) {
  if (ctx == nullptr) return "null context";
  return ctx->impl.last_error_.c_str();
}

int silk_hip_set_global(silk_hip_ctx* ctx, const silk_hip_global* g) {
  return guarded(ctx, [&](Context& c) {
    if (g == nullptr) return int{SILK_HIP_ERR_INVALID_CONFIG};
    const char* bad = nullptr;
    if (g->mas_precision < SILK_HIP_MAS_FP32 || g->mas_precision > SILK_HIP_MAS_BF16)
      bad = "mas_precision must be SILK_HIP_MAS_FP32, _FP16 or _BF16";
    else if (g->pcg_criterion != SILK_HIP_PCG_RR0 && g->pcg_criterion != SILK_HIP_PCG_RHS &&
             g->pcg_criterion != SILK_HIP_PCG_TRUE)
      bad = "pcg_criterion must be SILK_HIP_PCG_RR0, SILK_HIP_PCG_RHS or SILK_HIP_PCG_TRUE";
    else if (g->warm_start != SILK_HIP_WARM_PREVIOUS && g->warm_start != SILK_HIP_WARM_LEARNED &&
             g->warm_start != SILK_HIP_WARM_EXTRAPOLATE)
      bad = "warm_start must be SILK_HIP_WARM_PREVIOUS, SILK_HIP_WARM_LEARNED or "
            "SILK_HIP_WARM_EXTRAPOLATE";
    else if (g->warm_projection != SILK_HIP_PROJECT_RESIDUAL &&
             g->warm_projection != SILK_HIP_PROJECT_GALERKIN)
      bad = "warm_projection must be SILK_HIP_PROJECT_RESIDUAL or SILK_HIP_PROJECT_GALERKIN";
    else if (g->ccd_zero_toi != SILK_HIP_CCD_ZERO_SKIP &&
             g->ccd_zero_toi != SILK_HIP_CCD_ZERO_CONTACT)
      bad = "ccd_zero_toi must be SILK_HIP_CCD_ZERO_SKIP or SILK_HIP_CCD_ZERO_CONTACT";
    else if (g->admm_graph != SILK_HIP_GRAPH_OFF && g->admm_graph != SILK_HIP_GRAPH_ON)
      bad = "admm_graph must be SILK_HIP_GRAPH_OFF or SILK_HIP_GRAPH_ON";
    else if (g->warm_start == SILK_HIP_WARM_LEARNED && !c.has_warm_net())
      bad = "warm_start = SILK_HIP_WARM_LEARNED needs a network "
            "(silk_hip_set_warm_start_net)";
    if (bad != nullptr) {
      c.last_error_ = bad;
      return int{SILK_HIP_ERR_INVALID_CONFIG};
    }
    c.global_ = *g;
    return int{SILK_HIP_OK};
  });
}

int silk_hip_add_cloth(silk_hip_ctx* ctx, const silk_hip_cloth_params* cloth,
                       const silk_hip_collision_params* collision,
                       const float* verts, int64_t nv, const int32_t* faces,
                       int64_t nf, const int32_t* pins, int64_t npin,
                       uint32_t* handle) {
  return guarded(ctx, [&](Context& c) {
    if (handle == nullptr || cloth == nullptr || collision == nullptr)
      return int{SILK_HIP_ERR_INVALID_CONFIG};
    TriMesh mesh;
    std::string why;
    if (!silk::amd::make_cloth_mesh(verts, 3 * nv, faces, 3 * nf, mesh, why)) {
      *handle = 0;
      c.last_error_ = why;
      return int{SILK_HIP_ERR_INVALID_MESH};
    }
    if (!valid_pins(pins, npin, mesh.nv())) {
      c.last_error_ = "pin index out of range";
      return int{SILK_HIP_ERR_INVALID_PIN};
    }
    c.invalidate_scene();
    ClothEntity e;
    e.handle = c.new_handle();
    e.cloth = *cloth;
    e.coll = *collision;
    e.mesh = std::move(mesh);
    e.pins.assign(pins, pins + npin);
    c.cloths_.push_back(std::move(e));
    *handle = c.cloths_.back().handle;
    return int{SILK_HIP_OK};
  });
}

int silk_hip_add_obstacle(silk_hip_ctx* ctx,
                          const silk_hip_collision_params* collision,
                          const float* verts, int64_t nv,
                          const int32_t* faces, int64_t nf, uint32_t* handle) {
  return guarded(ctx, [&](Context& c) {
    if (handle == nullptr || collision == nullptr)
      return int{SILK_HIP_ERR_INVALID_CONFIG};
    TriMesh mesh;
    std::string why;
    if (!silk::amd::make_obstacle_mesh(verts, 3 * nv, faces, 3 * nf, mesh,
                                       why)) {
      *handle = 0;
      c.last_error_ = why;
      return int{SILK_HIP_ERR_INVALID_MESH};
    }
    c.invalidate_scene();
    ObstacleEntity o;
    o.handle = c.new_handle();
    o.coll = *collision;
    o.mesh = std::move(mesh);
    c.obstacles_.push_back(std::move(o));
    *handle = c.obstacles_.back().handle;
    return int{SILK_HIP_OK};
  });
}

int silk_hip_remove(silk_hip_ctx* ctx, uint32_t handle, int32_t is_cloth) {
  return guarded(ctx, [&](Context& c) {
    if (is_cloth) {
      if (!c.find_cloth(handle)) return int{SILK_HIP_ERR_INVALID_HANDLE};
      c.invalidate_scene();
      c.cloths_.erase(std::find_if(
          c.cloths_.begin(), c.cloths_.end(),
          [&](const ClothEntity& e) { return e.handle == handle; }));
    } else {
      if (!c.find_obstacle(handle)) return int{SILK_HIP_ERR_INVALID_HANDLE};
      c.invalidate_scene();
      c.obstacles_.erase(std::find_if(
          c.obstacles_.begin(), c.obstacles_.end(),
          [&](const ObstacleEntity& e) { return e.handle == handle; }));
    }
    return int{SILK_HIP_OK};
  });
}

int silk_hip_set_cloth_params(silk_hip_ctx* ctx, uint32_t handle,
                              const silk_hip_cloth_params* cloth) {
  return guarded(ctx, [&](Context& c) {
    ClothEntity* e = c.find_cloth(handle);
    if (!e) return int{SILK_HIP_ERR_INVALID_HANDLE};
    if (cloth == nullptr) return int{SILK_HIP_ERR_INVALID_CONFIG};
    c.invalidate_scene();
    e = c.find_cloth(handle);
    e->cloth = *cloth;
    return int{SILK_HIP_OK};
  });
}

int silk_hip_set_collision_params(silk_hip_ctx* ctx, uint32_t handle,
                                  int32_t is_cloth,
                                  const silk_hip_collision_params* coll) {
  return guarded(ctx, [&](Context& c) {
    if (coll == nullptr) return int{SILK_HIP_ERR_INVALID_CONFIG};
    if (is_cloth) {
      if (!c.find_cloth(handle)) return int{SILK_HIP_ERR_INVALID_HANDLE};
      c.invalidate_scene();
      c.find_cloth(handle)->coll = *coll;
    } else {
      if (!c.find_obstacle(handle)) return int{SILK_HIP_ERR_INVALID_HANDLE};
      c.invalidate_scene();
      c.find_obstacle(handle)->coll = *coll;
    }
    return int{SILK_HIP_OK};
  });
}

int silk_hip_set_pin_index(silk_hip_ctx* ctx, uint32_t handle,
                           const int32_t* pins, int64_t npin) {
  return guarded(ctx, [&](Context& c) {
    ClothEntity* e = c.find_cloth(handle);
    if (!e) return int{SILK_HIP_ERR_INVALID_HANDLE};
    for (int64_t i = 0; i < npin; ++i)
      if (pins == nullptr || pins[i] < 0 || pins[i] >= e->mesh.nv())
        return int{SILK_HIP_ERR_INVALID_PIN};
    c.invalidate_scene();
    e = c.find_cloth(handle);
    e->pins.assign(pins, pins + npin);
    e->has_pin_pos = false;  // re-seeded from the initial mesh
    return int{SILK_HIP_OK};
  });
}

int silk_hip_set_pin_positions(silk_hip_ctx* ctx, uint32_t handle,
                               const float* positions, int64_t n) {
  return guarded(ctx, [&](Context& c) {
    ClothEntity* e = c.find_cloth(handle);
    if (!e) return int{SILK_HIP_ERR_INVALID_HANDLE};
    if (n != 3 * static_cast<int64_t>(e->pins.size()) ||
        (n > 0 && positions == nullptr))
      return int{SILK_HIP_ERR_INVALID_PIN};
    e->pin_curr.assign(positions, positions + n);
    e->has_pin_pos = true;
    c.update_pin_targets(*e);
    return int{SILK_HIP_OK};
  });
}

int silk_hip_set_obstacle_positions(silk_hip_ctx* ctx, uint32_t handle,
                                    const float* positions, int64_t n) {
  return guarded(ctx, [&](Context& c) {
    ObstacleEntity* o = c.find_obstacle(handle);
    if (!o) return int{SILK_HIP_ERR_INVALID_HANDLE};
    if (n != static_cast<int64_t>(o->mesh.V.size()) || positions == nullptr)
      return int{SILK_HIP_ERR_INVALID_POSITION};
    if (o->has_pos) {
      o->is_static = false;
      o->is_static_twice = false;
      std::swap(o->curr, o->prev);
      o->curr.assign(positions, positions + n);
    } else {
      // A fresh record is static until the next change
      // (cuda_backend.cu:339-349, pin.cpp:337-345).
      o->curr.assign(positions, positions + n);
      o->prev = o->curr;
      o->is_static = true;
      o->is_static_twice = true;
      o->has_pos = true;
    }
    return int{SILK_HIP_OK};
  });
}

int silk_hip_step(silk_hip_ctx* ctx, silk_hip_step_stats* stats) {
  return guarded(ctx, [&](Context& c) { return c.step(stats); });
}

int silk_hip_reset(silk_hip_ctx* ctx) {
  return guarded(ctx, [&](Context& c) {
    c.cpu_world_.reset();
    c.scene_.reset();
    c.scene_dirty_ = true;
    for (auto& e : c.cloths_) {
      e.has_state = false;
      e.has_pin_pos = false;
    }
    for (auto& o : c.obstacles_) {
      o.has_pos = false;
      o.is_static = true;
      o.is_static_twice = true;
    }
    return int{SILK_HIP_OK};
  });
}

int silk_hip_get_positions(const silk_hip_ctx* cctx, uint32_t handle,
                           float* out, int64_t n) {
  silk_hip_ctx* ctx = const_cast<silk_hip_ctx*>(cctx);
  return guarded(ctx, [&](Context& c) {
    ClothEntity* e = c.find_cloth(handle);
    if (!e) return int{SILK_HIP_ERR_INVALID_HANDLE};
    if (out == nullptr || n != static_cast<int64_t>(e->mesh.V.size()))
      return int{SILK_HIP_ERR_INVALID_POSITION};
    c.get_positions(*e, out);
    return int{SILK_HIP_OK};
  });
}

int silk_hip_get_partition(const silk_hip_ctx* cctx, uint32_t handle,
                           int32_t* perm, int64_t nv, int32_t* offsets,
                           int64_t offsets_cap, int32_t* nparts) {
  silk_hip_ctx* ctx = const_cast<silk_hip_ctx*>(cctx);
  return guarded(ctx, [&](Context& c) {
    ClothEntity* e = c.find_cloth(handle);
    if (!e) return int{SILK_HIP_ERR_INVALID_HANDLE};
    if (!e->part_valid) {
      e->part = silk::amd::partition_vertices(e->mesh.V,
                                              silk::amd::MasTopology::kBank);
      e->pmesh = silk::amd::permute_mesh(e->mesh, e->part);
      e->part_valid = true;
      e->geom_valid = false;
    }
    const int np = static_cast<int>(e->part.offsets.size()) - 1;
    if (nparts) *nparts = np;
    if (perm != nullptr) {
      if (nv != e->mesh.nv()) return int{SILK_HIP_ERR_INVALID_POSITION};
      std::copy(e->part.perm.begin(), e->part.perm.end(), perm);
    }
    if (offsets != nullptr) {
      if (offsets_cap < np + 1) return int{SILK_HIP_ERR_INVALID_POSITION};
      std::copy(e->part.offsets.begin(), e->part.offsets.end(), offsets);
    }
    return int{SILK_HIP_OK};
  });
}

int silk_hip_get_timing(silk_hip_ctx* ctx, int32_t cls, double* ms,
                        int64_t* launches, int32_t reset) {
  return guarded(ctx, [&](Context& c) {
    if (cls < 0 || cls >= silk::amd::kTimeClasses) return int{SILK_HIP_ERR_INVALID_CONFIG};
    if (ms) *ms = c.timing_ms_[cls];
    if (launches) *launches = c.timing_launch_[cls];
    if (reset) {
      c.timing_ms_[cls] = 0.0;
      c.timing_launch_[cls] = 0;
    }
    return int{SILK_HIP_OK};
  });
}

int silk_hip_set_timing(silk_hip_ctx* ctx, int32_t on) {
  return guarded(ctx, [&](Context& c) {
    if (on < 0) return int{SILK_HIP_ERR_INVALID_CONFIG};
    // 1 times every class; any other value is a mask of classes (bit c).
    c.timing_mask_ = on == 1 ? ~0u : static_cast<uint32_t>(on);
    return int{SILK_HIP_OK};
  });
}

int silk_hip_get_scene_info(silk_hip_ctx* ctx, int64_t* info, int32_t n) {
  return guarded(ctx, [&](Context& c) {
    if (info == nullptr || n < 0) return int{SILK_HIP_ERR_INVALID_CONFIG};
    if (c.cloths_.empty()) return int{SILK_HIP_ERR_NEED_INIT_SOLVER};
    if (c.scene_dirty_ || !c.scene_) c.build_scene();
    const auto& S = *c.scene_;
    int ncb = 0;
    for (int b = 0; b < S.nbank; ++b) ncb += S.h_bank_level[b] > 0;
    const int64_t v[12] = {S.nv, S.nf, S.nnz_off, S.nbank,
                           S.pers.nslot_total,
                           c.persistent_on() ? S.pers.grid : 0, ncb,
                           S.coll ? S.coll->total_vertices() : 0,
                           c.persistent_on() ? static_cast<int64_t>(S.ps_batches.size()) : 0,
                           S.nsys,
                           static_cast<int64_t>(silk::amd::pcg_persistent_lds(S.pers)),
                           S.pers.dense};
    for (int i = 0; i < n && i < 12; ++i) info[i] = v[i];
    return int{SILK_HIP_OK};
  });
}

int silk_hip_collider_vertex_count(silk_hip_ctx* ctx, int64_t* nvt) {
  return guarded(ctx, [&](Context& c) {
    if (nvt == nullptr) return int{SILK_HIP_ERR_INVALID_CONFIG};
    if (c.cloths_.empty()) {
      *nvt = 0;
      return int{SILK_HIP_OK};
    }
    if (c.scene_dirty_ || !c.scene_) c.build_scene();
    *nvt = c.scene_->coll ? c.scene_->coll->total_vertices() : 0;
    return int{SILK_HIP_OK};
  });
}

int silk_hip_debug_candidates(silk_hip_ctx* ctx, const float* pos0,
                              const float* pos1, int64_t n, int32_t* pt,
                              int64_t pt_cap, int64_t* npt, int32_t* ee,
                              int64_t ee_cap, int64_t* nee) {
  return guarded(ctx, [&](Context& c) {
    if (npt == nullptr || nee == nullptr) return int{SILK_HIP_ERR_INVALID_CONFIG};
    if (c.cloths_.empty()) return int{SILK_HIP_ERR_NEED_INIT_SOLVER};
    if (c.scene_dirty_ || !c.scene_) c.build_scene();
    auto* cw = c.scene_->coll.get();
    if (cw == nullptr) {
      *npt = 0;
      *nee = 0;
      return int{SILK_HIP_OK};
    }
    if (pos0 == nullptr || pos1 == nullptr ||
        n != 3 * static_cast<int64_t>(cw->total_vertices()))
      return int{SILK_HIP_ERR_INVALID_POSITION};
    std::vector<int> vpt, vee;
    cw->debug_candidates(pos0, pos1, vpt, vee);
    *npt = static_cast<int64_t>(vpt.size() / 4);
    *nee = static_cast<int64_t>(vee.size() / 4);
    if (pt != nullptr)
      std::copy_n(vpt.begin(), 4 * std::min<int64_t>(pt_cap, *npt), pt);
    if (ee != nullptr)
      std::copy_n(vee.begin(), 4 * std::min<int64_t>(ee_cap, *nee), ee);
    return int{SILK_HIP_OK};
  });
}

int silk_hip_debug_dcd(silk_hip_ctx* ctx, const float* pos0, const float* pos1,
                       int64_t n, float time, float dhat, int32_t* recs,
                       float* dist, int64_t cap, int64_t* nrec) {
  return guarded(ctx, [&](Context& c) {
    if (nrec == nullptr || !(time >= 0.0f && time <= 1.0f) || !(dhat > 0.0f))
      return int{SILK_HIP_ERR_INVALID_CONFIG};
    if (c.cloths_.empty()) return int{SILK_HIP_ERR_NEED_INIT_SOLVER};
    if (c.scene_dirty_ || !c.scene_) c.build_scene();
    auto* cw = c.scene_->coll.get();
    if (cw == nullptr) {
      *nrec = 0;
      return int{SILK_HIP_OK};
    }
    if (pos0 == nullptr || pos1 == nullptr ||
        n != 3 * static_cast<int64_t>(cw->total_vertices()))
      return int{SILK_HIP_ERR_INVALID_POSITION};
    std::vector<int> vr;
    std::vector<float> vd;
    cw->debug_dcd(pos0, pos1, time, dhat, vr, vd);
    *nrec = static_cast<int64_t>(vd.size());
    const int64_t m = std::min<int64_t>(cap, *nrec);
    if (recs != nullptr) std::copy_n(vr.begin(), 6 * m, recs);
    if (dist != nullptr) std::copy_n(vd.begin(), m, dist);
    return int{SILK_HIP_OK};
  });
}

int silk_hip_debug_decisions(silk_hip_ctx* ctx, float* tois, int32_t* admm_iters,
                             int32_t cap_toi, int32_t* n_toi, int32_t* contacts,
                             int32_t* contact_info, int64_t cap_contacts,
                             int64_t* n_contacts) {
  return guarded(ctx, [&](Context& c) {
    if (n_toi == nullptr || n_contacts == nullptr) return int{SILK_HIP_ERR_INVALID_CONFIG};
    const std::vector<float>& tl = c.toi_log_;
    *n_toi = static_cast<int32_t>(tl.size());
    const int64_t nt = std::min<int64_t>(cap_toi, *n_toi);
    if (tois != nullptr) std::copy_n(tl.begin(), nt, tois);
    if (admm_iters != nullptr) std::copy_n(c.admm_log_.begin(), nt, admm_iters);
    const CollisionWorld* cw = c.scene_ ? c.scene_->coll.get() : nullptr;
    static const std::vector<int> none;
    const std::vector<int>& log = cw != nullptr ? cw->decision_log() : none;
    const std::vector<int>& info = cw != nullptr ? cw->decision_info() : none;
    *n_contacts = static_cast<int64_t>(log.size() / 6);
    const int64_t nc = std::min<int64_t>(cap_contacts, *n_contacts);
    if (contacts != nullptr) std::copy_n(log.begin(), 6 * nc, contacts);
    if (contact_info != nullptr) std::copy_n(info.begin(), nc, contact_info);
    return int{SILK_HIP_OK};
  });
}

int silk_hip_debug_ccd(silk_hip_ctx* ctx, int64_t n, const int32_t* is_vf,
                       const double* p0, const double* p1, const double* err,
                       double ms, double tolerance, int32_t max_iter,
                       int32_t* hit, double* toi) {
  return guarded(ctx, [&](Context& c) {
    if (n < 0 || n > (1 << 24)) return int{SILK_HIP_ERR_INVALID_CONFIG};
    if (n == 0) return int{SILK_HIP_OK};
    if (!is_vf || !p0 || !p1 || !err || !hit || !toi)
      return int{SILK_HIP_ERR_INVALID_CONFIG};
    c.ensure_device();
    using silk::amd::DevBuf;
    const size_t m = static_cast<size_t>(n);
    std::vector<int> hv(is_vf, is_vf + m);
    std::vector<double> h0(p0, p0 + 12 * m), h1(p1, p1 + 12 * m),
        he(err, err + 3 * m);
    DevBuf<int> dv, dhit;
    DevBuf<double> d0, d1, de, dtoi;
    dv.upload(hv, c.stream_);
    d0.upload(h0, c.stream_);
    d1.upload(h1, c.stream_);
    de.upload(he, c.stream_);
    dhit.resize(m);
    dtoi.resize(m);
    silk::amd::launch_ccd_each(c.stream_, static_cast<int>(n), dv.get(),
                               d0.get(), d1.get(), de.get(), ms, tolerance,
                               max_iter, dhit.get(), dtoi.get(), c.ccd_zero_mode());
    silk::amd::hip_check(hipGetLastError(), "ccd launch");
    silk::amd::hip_check(hipMemcpyAsync(hit, dhit.get(), m * sizeof(int),
                                        hipMemcpyDeviceToHost, c.stream_),
                         "ccd hit");
    silk::amd::hip_check(hipMemcpyAsync(toi, dtoi.get(), m * sizeof(double),
                                        hipMemcpyDeviceToHost, c.stream_),
                         "ccd toi");
    silk::amd::hip_check(hipStreamSynchronize(c.stream_), "ccd sync");
    return int{SILK_HIP_OK};
  });
}

int silk_hip_debug_ccd_min(silk_hip_ctx* ctx, int64_t n, const int32_t* is_vf,
                           const double* p0, const double* p1,
                           const double* err, double ms, double tolerance,
                           int32_t max_iter, double* min_toi) {
  return guarded(ctx, [&](Context& c) {
    if (n < 0 || n > (1 << 24)) return int{SILK_HIP_ERR_INVALID_CONFIG};
    if (min_toi == nullptr) return int{SILK_HIP_ERR_INVALID_CONFIG};
    *min_toi = 1.0;
    if (n == 0) return int{SILK_HIP_OK};
    if (!is_vf || !p0 || !p1 || !err) return int{SILK_HIP_ERR_INVALID_CONFIG};
    c.ensure_device();
    using silk::amd::DevBuf;
    const size_t m = static_cast<size_t>(n);
    std::vector<int> hv(is_vf, is_vf + m);
    std::vector<double> h0(p0, p0 + 12 * m), h1(p1, p1 + 12 * m),
        he(err, err + 3 * m), one{1.0};
    DevBuf<int> dv;
    DevBuf<double> d0, d1, de, dmin;
    dv.upload(hv, c.stream_);
    d0.upload(h0, c.stream_);
    d1.upload(h1, c.stream_);
    de.upload(he, c.stream_);
    dmin.upload(one, c.stream_);
    silk::amd::launch_ccd_min(
        c.stream_, static_cast<int>(n), dv.get(), d0.get(), d1.get(), de.get(),
        ms, tolerance, max_iter,
        reinterpret_cast<unsigned long long*>(dmin.get()), c.ccd_zero_mode());
    silk::amd::hip_check(hipGetLastError(), "ccd launch");
    silk::amd::hip_check(hipMemcpyAsync(min_toi, dmin.get(), sizeof(double),
                                        hipMemcpyDeviceToHost, c.stream_),
                         "ccd min");
    silk::amd::hip_check(hipStreamSynchronize(c.stream_), "ccd sync");
    return int{SILK_HIP_OK};
  });
}

int silk_hip_ccd_min_noalloc(silk_hip_ctx* ctx, int64_t n, const int32_t* is_vf,
                             const double* p0, const double* p1, const double* err,
                             double ms, double tolerance, int32_t max_iter,
                             double* min_toi) {
  return guarded(ctx, [&](Context& c) {
    return silk::amd::ccd_min_noalloc(c, n, is_vf, p0, p1, err, ms, tolerance, max_iter,
                                      min_toi);
  });
}

int silk_hip_debug_coulomb(silk_hip_ctx* ctx, int64_t n, const float* u,
                           const float* normal, const float* mu, float* out) {
  return guarded(ctx, [&](Context& c) {
    if (n < 0 || n > (1 << 24)) return int{SILK_HIP_ERR_INVALID_CONFIG};
    if (n == 0) return int{SILK_HIP_OK};
    if (!u || !normal || !mu || !out) return int{SILK_HIP_ERR_INVALID_CONFIG};
    c.ensure_device();
    using silk::amd::DevBuf;
    const size_t m = static_cast<size_t>(n);
    std::vector<float> hu(u, u + 3 * m), hn(normal, normal + 3 * m), hm(mu, mu + m),
        ho(3 * m, 0.0f);
    DevBuf<float> du, dn, dm, dout;
    du.upload(hu, c.stream_);
    dn.upload(hn, c.stream_);
    dm.upload(hm, c.stream_);
    dout.upload(ho, c.stream_);
    silk::amd::launch_coulomb_each(c.stream_, static_cast<int>(n), du.get(), dn.get(),
                                   dm.get(), dout.get());
    silk::amd::hip_check(hipGetLastError(), "coulomb launch");
    silk::amd::hip_check(hipMemcpyAsync(out, dout.get(), 3 * m * sizeof(float),
                                        hipMemcpyDeviceToHost, c.stream_),
                         "coulomb out");
    silk::amd::hip_check(hipStreamSynchronize(c.stream_), "coulomb sync");
    return int{SILK_HIP_OK};
  });
}

int silk_hip_debug_copy(silk_hip_ctx* ctx, int64_t bytes, int32_t reps, double* ms) {
  return guarded(ctx, [&](Context& c) {
    if (bytes < 4 || bytes > (int64_t{1} << 32) || reps < 1 || reps > 100)
      return int{SILK_HIP_ERR_INVALID_CONFIG};
    c.ensure_device();
    using silk::amd::DevBuf;
    const int64_t n = bytes / 4;
    DevBuf<float> src, dst;
    src.resize(static_cast<size_t>(n));
    dst.resize(static_cast<size_t>(n));
    silk::amd::hip_check(hipMemsetAsync(src.get(), 0, static_cast<size_t>(n) * 4, c.stream_),
                         "copy init");
    hipEvent_t e0, e1;
    silk::amd::hip_check(hipEventCreate(&e0), "event");
    silk::amd::hip_check(hipEventCreate(&e1), "event");
    silk::amd::hip_check(hipEventRecord(e0, c.stream_), "event");
    for (int r = 0; r < reps; ++r)
      silk::amd::launch_copy_unless(c.stream_, dst.get(), src.get(), n, nullptr);
    silk::amd::hip_check(hipEventRecord(e1, c.stream_), "event");
    silk::amd::hip_check(hipStreamSynchronize(c.stream_), "copy sync");
    float el = 0.0f;
    silk::amd::hip_check(hipEventElapsedTime(&el, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms != nullptr) *ms = el;
    return int{SILK_HIP_OK};
  });
}

int silk_hip_debug_pcg(silk_hip_ctx* ctx, const float* d, int64_t nv,
                       const float* b, int64_t n3, float rel_tol, float* x,
                       int32_t* iterations) {
  return guarded(ctx, [&](Context& c) {
    if (!d || !b || !x || !iterations || !(rel_tol >= 0.0f))
      return int{SILK_HIP_ERR_INVALID_CONFIG};
    if (c.cloths_.empty()) return int{SILK_HIP_ERR_INVALID_CONFIG};
    int64_t total = 0;
    for (const auto& e : c.cloths_) total += e.mesh.nv();
    if (nv != total || n3 != 3 * total) return int{SILK_HIP_ERR_INVALID_CONFIG};
    *iterations = c.debug_pcg(d, b, rel_tol, x);
    return int{SILK_HIP_OK};
  });
}

int silk_hip_autotune(silk_hip_ctx* ctx, int32_t* report, int32_t cap) {
  return guarded(ctx, [&](Context& c) {
    if (report == nullptr || cap < SILK_HIP_AUTOTUNE_REPORT)
      return int{SILK_HIP_ERR_INVALID_CONFIG};
    return c.autotune(report);
  });
}

int silk_hip_set_warm_start_net(silk_hip_ctx* ctx, const float* params, int64_t n) {
  return guarded(ctx, [&](Context& c) {
    if (params == nullptr || n != SILK_HIP_WARM_NET_PARAMS)
      return int{SILK_HIP_ERR_INVALID_CONFIG};
    for (int64_t i = 0; i < n; ++i)
      if (!std::isfinite(params[i])) {
        c.last_error_ = "non-finite network parameter";
        return int{SILK_HIP_ERR_INVALID_CONFIG};
      }
    c.warm_set(params);
    return int{SILK_HIP_OK};
  });
}

int silk_hip_get_warm_start_net(silk_hip_ctx* ctx, float* params, int64_t n) {
  return guarded(ctx, [&](Context& c) {
    if (params == nullptr || n != SILK_HIP_WARM_NET_PARAMS)
      return int{SILK_HIP_ERR_INVALID_CONFIG};
    if (!c.has_warm_net()) return int{SILK_HIP_ERR_NEED_INIT_SOLVER};
    c.warm_get(params);
    return int{SILK_HIP_OK};
  });
}

int silk_hip_warm_start_record(silk_hip_ctx* ctx, int64_t capacity, float rate) {
  return guarded(ctx, [&](Context& c) {
    if (capacity < 0 || capacity > (int64_t{1} << 31) || !(rate >= 0.0f && rate <= 1.0f))
      return int{SILK_HIP_ERR_INVALID_CONFIG};
    c.warm_record(capacity, rate);
    return int{SILK_HIP_OK};
  });
}

int silk_hip_warm_start_samples(silk_hip_ctx* ctx, int64_t* n) {
  return guarded(ctx, [&](Context& c) {
    if (n == nullptr) return int{SILK_HIP_ERR_INVALID_CONFIG};
    *n = c.recording() ? c.warm_samples() : 0;
    return int{SILK_HIP_OK};
  });
}

int silk_hip_warm_start_train(silk_hip_ctx* ctx, int32_t steps, int32_t batch,
                              float lr, uint64_t seed, float* loss) {
  return guarded(ctx, [&](Context& c) {
    if (steps < 0 || batch <= 0 || batch > (1 << 24) || !(lr > 0.0f))
      return int{SILK_HIP_ERR_INVALID_CONFIG};
    if (!c.recording() || c.warm_samples() == 0) {
      c.last_error_ = "no recorded samples (silk_hip_warm_start_record)";
      return int{SILK_HIP_ERR_NEED_INIT_SOLVER};
    }
    c.warm_train(steps, batch, lr, seed, loss);
    return int{SILK_HIP_OK};
  });
}

int silk_hip_debug_warm_net(silk_hip_ctx* ctx, int64_t n, const float* feat,
                            const float* target, float* y, float* grad,
                            double* loss) {
  return guarded(ctx, [&](Context& c) {
    if (n <= 0 || n > (1 << 22) || !feat || !target || !y || !grad || !loss)
      return int{SILK_HIP_ERR_INVALID_CONFIG};
    if (!c.has_warm_net()) return int{SILK_HIP_ERR_NEED_INIT_SOLVER};
    c.warm_debug(n, feat, target, y, grad, loss);
    return int{SILK_HIP_OK};
  });
}

}  // extern "C"
