// Runtime context: registry, device scene assembly and the PD step.
#include "api/context.hpp"

#include "cpu/cpu_world.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <numeric>

#include "api/collision.hpp"

namespace silk::amd {

namespace {

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

template <typename T>
void download(std::vector<T>& h, const T* d, size_t n, hipStream_t s) {
  h.resize(n);
  if (n > 0)
    hip_check(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost,
                             s),
              "download");
}

}  // namespace

Scene::Scene() {
  hip_check(hipHostMalloc(reinterpret_cast<void**>(&h_admm_ctl),
                          2 * sizeof(AdmmCtl)),
            "hipHostMalloc");
  std::memset(static_cast<void*>(h_admm_ctl), 0, 2 * sizeof(AdmmCtl));
}

Scene::~Scene() {
  if (admm_exec != nullptr) (void)hipGraphExecDestroy(admm_exec);
  if (h_admm_ctl != nullptr) (void)hipHostFree(h_admm_ctl);
}

Context::Context(int device) : device_(device), cpu_(device == SILK_HIP_DEVICE_CPU) {
  global_.acceleration[0] = 0.0f;
  global_.acceleration[1] = 0.0f;
  global_.acceleration[2] = -10.0f;
  global_.max_outer_iteration = 100;
  global_.max_inner_iteration = 100;
  global_.dt = 1.0f / 60.0f;
  global_.linear_solver_abs_tol = 1e-7f;
  global_.linear_solver_rel_tol_min = 1e-6f;
  global_.linear_solver_rel_tol_max = 1e-3f;
  global_.initial_linear_rel_tol = 1e-2f;
  global_.admm_abs_tol = 1e-5f;
  global_.admm_rel_tol = 1e-3f;
}

void Context::timed_begin(int cls) {
  timed_open_ = false;
  if (!timing(cls)) return;
  const size_t k = ev_cls_.size();
  if (k >= ev_pool_.size()) {
    EvPair e;
    hip_check(hipEventCreate(&e.a), "hipEventCreate");
    hip_check(hipEventCreate(&e.b), "hipEventCreate");
    ev_pool_.push_back(e);
  }
  ev_cls_.push_back(cls);
  hip_check(hipEventRecord(ev_pool_[k].a, stream_), "event");
  timed_open_ = true;
}

void Context::timed_end() {
  if (!timed_open_) return;
  timed_open_ = false;
  hip_check(hipEventRecord(ev_pool_[ev_cls_.size() - 1].b, stream_), "event");
}

void Context::timed_collect() {
  for (size_t k = 0; k < ev_cls_.size(); ++k) {
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, ev_pool_[k].a, ev_pool_[k].b), "elapsed");
    timing_ms_[ev_cls_[k]] += ms;
    timing_launch_[ev_cls_[k]] += 1;
  }
  ev_cls_.clear();
}

Context::~Context() {
  scene_.reset();
  for (auto& e : ev_pool_) {
    if (e.a) (void)hipEventDestroy(e.a);
    if (e.b) (void)hipEventDestroy(e.b);
  }
  if (dev_ready_) {
    for (auto& e : ev_)
      if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
  }
}

uint32_t Context::new_handle() {
  for (;;) {
    uint32_t h = static_cast<uint32_t>(splitmix64(rng_state_));
    if (h == 0) continue;
    if (find_cloth(h) || find_obstacle(h)) continue;
    return h;
  }
}

ClothEntity* Context::find_cloth(uint32_t h) {
  for (auto& c : cloths_)
    if (c.handle == h) return &c;
  return nullptr;
}

ObstacleEntity* Context::find_obstacle(uint32_t h) {
  for (auto& o : obstacles_)
    if (o.handle == h) return &o;
  return nullptr;
}

void Context::ensure_device() {
  if (dev_ready_) return;
  if (cpu_) throw std::runtime_error("not available on the CPU backend");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device_ || device_ < 0) {
    (void)hipGetLastError();
    throw std::runtime_error("no HIP device");
  }
  hip_check(hipSetDevice(device_), "hipSetDevice");
  hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking),
            "hipStreamCreate");
  for (auto& e : ev_) hip_check(hipEventCreate(&e), "hipEventCreate");
  dev_ready_ = true;
}

void Context::invalidate_scene() {
  if (cpu_world_) {
    for (size_t ci = 0; ci < cloths_.size(); ++ci) {
      cloths_[ci].x = cpu_world_->x(static_cast<int>(ci));
      cloths_[ci].v = cpu_world_->v(static_cast<int>(ci));
      cloths_[ci].has_state = true;
    }
    cpu_world_.reset();
  }
  if (scene_ && dev_ready_) {
    // Persist cloth states (solver order on device -> caller order on host).
    std::vector<float> hx, hv;
    download(hx, scene_->x.get(), 3 * static_cast<size_t>(scene_->nv), stream_);
    download(hv, scene_->v.get(), 3 * static_cast<size_t>(scene_->nv), stream_);
    hip_check(hipStreamSynchronize(stream_), "sync");
    for (size_t ci = 0; ci < cloths_.size(); ++ci) {
      ClothEntity& c = cloths_[ci];
      if (ci + 1 >= scene_->vert_off.size()) break;
      const int off = scene_->vert_off[ci];
      const int nv = c.mesh.nv();
      c.x.assign(3 * nv, 0.0f);
      c.v.assign(3 * nv, 0.0f);
      for (int v = 0; v < nv; ++v)
        for (int d = 0; d < 3; ++d) {
          c.x[3 * v + d] = hx[3 * (off + c.part.perm[v]) + d];
          c.v[3 * v + d] = hv[3 * (off + c.part.perm[v]) + d];
        }
      c.has_state = true;
    }
  }
  scene_.reset();
  scene_dirty_ = true;
}

void Context::ensure_cloth_derived(ClothEntity& c) {
  if (!c.part_valid) {
    c.part = partition_vertices(c.mesh.V, MasTopology::kBank);
    c.pmesh = permute_mesh(c.mesh, c.part);
    c.part_valid = true;
    c.geom_valid = false;
  }
  if (!c.geom_valid) {
    c.geom = build_cloth_geometry(c.pmesh);
    c.geom_valid = true;
  }
  if (!c.has_pin_pos) {
    c.pin_curr.assign(3 * c.pins.size(), 0.0f);
    for (size_t i = 0; i < c.pins.size(); ++i)
      for (int d = 0; d < 3; ++d)
        c.pin_curr[3 * i + d] = c.mesh.V[3 * c.pins[i] + d];
    c.has_pin_pos = true;
  }
  if (!c.has_state) {
    c.x = c.mesh.V;
    c.v.assign(c.mesh.V.size(), 0.0f);
    c.has_state = true;
  }
}

void Context::build_cpu_world() {
  invalidate_scene();
  int threads = 0;
  if (const char* e = std::getenv("SILK_CPU_THREADS")) threads = std::atoi(e);
  auto w = std::make_unique<cpu::CpuWorld>(global_, threads);
  for (ClothEntity& c : cloths_) {
    cpu::ClothInput in;
    in.mesh = &c.mesh;
    in.cloth = c.cloth;
    in.coll = c.coll;
    in.pins = c.pins;
    if (c.has_pin_pos) {
      in.pin_pos = c.pin_curr;
    } else {
      for (int p : c.pins)
        for (int d = 0; d < 3; ++d) in.pin_pos.push_back(c.mesh.V[3 * p + d]);
    }
    in.x = c.has_state ? c.x : c.mesh.V;
    in.v = c.has_state ? c.v : std::vector<float>(c.mesh.V.size(), 0.0f);
    w->add_cloth(in);
  }
  for (ObstacleEntity& o : obstacles_) {
    cpu::ObstacleInput in;
    in.mesh = &o.mesh;
    in.coll = o.coll;
    w->add_obstacle(in);
  }
  cpu_world_ = std::move(w);
  cpu_global_ = global_;
  scene_dirty_ = false;
}

int Context::cpu_step(silk_hip_step_stats* stats) {
  if (!cpu_world_ || scene_dirty_ ||
      std::memcmp(&cpu_global_, &global_, sizeof(global_)) != 0) {
    try {
      build_cpu_world();
    } catch (const std::runtime_error& e) {
      last_error_ = e.what();
      return SILK_HIP_ERR_CHOLESKY_FAIL;
    }
  }
  for (size_t i = 0; i < obstacles_.size(); ++i) {
    const ObstacleEntity& o = obstacles_[i];
    const std::vector<float>& prev = o.has_pos ? o.prev : o.mesh.V;
    const std::vector<float>& curr = o.has_pos ? o.curr : o.mesh.V;
    cpu_world_->set_obstacle_frame(static_cast<int>(i), prev.data(), curr.data());
  }
  cpu::StepStats st;
  if (!cpu_world_->step(&st)) {
    last_error_ = "solver diverged";
    return SILK_HIP_ERR_UNKNOWN;
  }
  for (auto& o : obstacles_) {
    if (o.has_pos) o.prev = o.curr;
    if (o.is_static) o.is_static_twice = true;
    else o.is_static = true;
  }
  if (stats) {
    stats->outer_iterations = st.outer;
    stats->admm_iterations = st.inner;
    stats->contacts = st.collisions;
    stats->ccd_candidates = st.ccd_candidates;
    stats->min_toi = st.min_toi;
  }
  return SILK_HIP_OK;
}

void Context::build_scene() {
  ensure_device();
  auto sc = std::make_unique<Scene>();
  Scene& S = *sc;
  hipStream_t st = stream_;
  const float h = global_.dt;
  S.dt = h;
  S.nsys = static_cast<int>(cloths_.size());
  S.vert_off.assign(S.nsys + 1, 0);
  S.face_off.assign(S.nsys + 1, 0);
  for (int ci = 0; ci < S.nsys; ++ci) {
    ensure_cloth_derived(cloths_[ci]);
    S.vert_off[ci + 1] = S.vert_off[ci] + cloths_[ci].mesh.nv();
    S.face_off[ci + 1] = S.face_off[ci] + cloths_[ci].mesh.nf();
  }
  S.nv = S.vert_off[S.nsys];
  S.nf = S.face_off[S.nsys];
  const int nv = S.nv, nf = S.nf;

  std::vector<float> hx(3 * nv), hv(3 * nv), hfparam(4 * nf), hfstiff(2 * nf),
      hmass(nv), hdamp(nv), hpin_t(3 * nv, 0.0f), hcw(nv);
  std::vector<double> hcw_d(nv);
  std::vector<double> hbend(3 * nv);
  std::vector<int> hF(3 * nf), hpinned(nv, 0);
  std::vector<std::vector<int>> vf(nv);

  // PCG / MAS host tables.
  std::vector<int> blk_sys, blk_row0, blk_rows, sys_blk0{0};
  std::vector<int> slice_off, slice_width, ell_col;
  std::vector<float> ell_val;
  std::vector<int> bank_sys, bank_level, bank_member;
  std::vector<float> bank_static;
  // coarse nodes: level-1 nodes of all systems first, then deeper ones.
  struct SysMas {
    MasTopology topo;
    int c1_off = 0;             // global id of first level-1 node
    std::vector<int> deep_off;  // per matrix level >= 2
  };
  std::vector<SysMas> smas(S.nsys);
  int n1 = 0;
  std::vector<CsrMatrix> Ks(S.nsys);

  for (int ci = 0; ci < S.nsys; ++ci) {
    ClothEntity& c = cloths_[ci];
    const int vo = S.vert_off[ci], fo = S.face_off[ci];
    const int cnv = c.mesh.nv(), cnf = c.mesh.nf();
    const double k = c.cloth.elastic_stiffness, kb = c.cloth.bending_stiffness;
    const double dens = c.cloth.density;
    // state (caller order -> solver order)
    for (int v = 0; v < cnv; ++v)
      for (int d = 0; d < 3; ++d) {
        hx[3 * (vo + c.part.perm[v]) + d] = c.x[3 * v + d];
        hv[3 * (vo + c.part.perm[v]) + d] = c.v[3 * v + d];
      }
    for (int f = 0; f < cnf; ++f) {
      for (int j = 0; j < 3; ++j) {
        const int lv = c.pmesh.F[3 * f + j];
        hF[3 * (fo + f) + j] = vo + lv;
        vf[vo + lv].push_back(4 * (fo + f) + j);
      }
      for (int j = 0; j < 4; ++j)
        hfparam[4 * (fo + f) + j] = c.geom.face_param[4 * f + j];
      hfstiff[2 * (fo + f)] = c.cloth.elastic_stiffness;
      hfstiff[2 * (fo + f) + 1] = c.cloth.elastic_stiffness;  // ADMM penalty
    }
    Ks[ci] = combine_operator(c.geom, k, kb);
    const CsrMatrix& K = Ks[ci];
    for (int v = 0; v < cnv; ++v) {
      const double m = dens * c.geom.mass[v];
      hmass[vo + v] = static_cast<float>(m);
      hdamp[vo + v] = 1.0f - c.cloth.damping;
      for (int d = 0; d < 3; ++d)
        hbend[3 * (vo + v) + d] = kb * c.geom.bending_rest[3 * v + d];
      // Daviet contact weight (cloth_assembly_l1_cache.cu:303-321): the 3x3
      // diagonal block is (K_vv + m/h^2) I3, so its smallest eigenvalue is
      // that scalar.
      double kvv = 0.0;
      for (int a = K.row_ptr[v]; a < K.row_ptr[v + 1]; ++a)
        if (K.col[a] == v) kvv = K.val[a];
      const double inertia = m / (static_cast<double>(h) * h);
      const double eta = kvv + inertia;
      hcw_d[vo + v] = std::max(0.001 * eta, std::min(25.0 * inertia, eta));
      hcw[vo + v] = static_cast<float>(hcw_d[vo + v]);
    }
    for (size_t i = 0; i < c.pins.size(); ++i) {
      const int lv = c.part.perm[c.pins[i]];
      if (!hpinned[vo + lv]) S.pinned_dof += 3;
      hpinned[vo + lv] = 1;
      for (int d = 0; d < 3; ++d)
        hpin_t[3 * (vo + lv) + d] = c.pin_curr[3 * i + d];
    }

    // Row blocks and sliced ELL of the off-diagonal entries (the operator is
    // applied in difference form, which needs no diagonal).
    auto offdiag = [&K](int row) {
      int cnt = 0;
      for (int a = K.row_ptr[row]; a < K.row_ptr[row + 1]; ++a)
        cnt += K.col[a] != row;
      return cnt;
    };
    for (int r0 = 0; r0 < cnv; r0 += kBlock) {
      const int rows = std::min(kBlock, cnv - r0);
      blk_sys.push_back(ci);
      blk_row0.push_back(vo + r0);
      blk_rows.push_back(rows);
      for (int sl = 0; sl < kBlock / kSlice; ++sl) {
        const int s0 = r0 + sl * kSlice;
        const int srows = std::max(0, std::min(kSlice, cnv - s0));
        int width = 0;
        for (int rr = 0; rr < srows; ++rr) width = std::max(width, offdiag(s0 + rr));
        slice_off.push_back(static_cast<int>(ell_col.size()));
        slice_width.push_back(width);
        const size_t base = ell_col.size();
        ell_col.resize(base + static_cast<size_t>(width) * kSlice);
        ell_val.resize(base + static_cast<size_t>(width) * kSlice);
        for (int lane = 0; lane < kSlice; ++lane) {
          const int row = s0 + lane;
          const bool real = lane < srows;
          int kk = 0;
          if (real) {
            for (int a = K.row_ptr[row]; a < K.row_ptr[row + 1]; ++a) {
              if (K.col[a] == row) continue;
              const size_t e = base + static_cast<size_t>(kk) * kSlice + lane;
              ell_col[e] = vo + K.col[a];
              ell_val[e] = static_cast<float>(K.val[a]);
              ++kk;
            }
          }
          for (; kk < width; ++kk) {
            const size_t e = base + static_cast<size_t>(kk) * kSlice + lane;
            ell_col[e] = real ? vo + row : vo;
            ell_val[e] = 0.0f;
          }
        }
      }
    }
    sys_blk0.push_back(static_cast<int>(blk_sys.size()));

    // MAS hierarchy.
    SysMas& sm = smas[ci];
    sm.topo = build_mas_topology(pattern_graph(K), c.part.offsets);
    sm.c1_off = n1;
    n1 += sm.topo.coarse_levels > 0 ? sm.topo.level_nodes[1] : 0;
  }
  // deep-node offsets
  int ncn = n1;
  for (int ci = 0; ci < S.nsys; ++ci) {
    SysMas& sm = smas[ci];
    for (int l = 2; l <= sm.topo.coarse_levels; ++l) {
      sm.deep_off.push_back(ncn);
      ncn += sm.topo.level_nodes[l];
    }
  }
  S.n1node = n1;
  S.ncnode = ncn;
  std::vector<int> cnode_ptr(ncn + 1, 0), cnode_level(ncn, 0),
      row_cnode(4 * static_cast<size_t>(nv), -1);
  // Persistent-PCG tables, slot order (banks padded to 32 slots).
  std::vector<PcgSysDev> ps_sys;
  std::vector<int> ps_slot_row, ps_slot_cn, ps_cb_bank, ps_cb_member,
      ps_cn_ptr{0}, ps_cn_list, ps_slot_of_row(nv, -1);
  int ps_wg = 0;
  std::vector<std::vector<int>> cmembers(ncn);
  for (int ci = 0; ci < S.nsys; ++ci) {
    const SysMas& sm = smas[ci];
    const MasTopology& t = sm.topo;
    const int vo = S.vert_off[ci];
    auto gid = [&](int mlevel, int local) {
      return mlevel == 1 ? sm.c1_off + local : sm.deep_off[mlevel - 2] + local;
    };
    for (int row = 0; row < t.n; ++row) {
      for (int l = 1; l <= t.coarse_levels; ++l) {
        const int g = gid(l, t.cmap[l - 1][row]);
        row_cnode[4 * static_cast<size_t>(vo + row) + (l - 1)] = g;
        cnode_level[g] = l;
        if (l == 1) cmembers[g].push_back(vo + row);
      }
    }
    // deep nodes list their level-1 descendants (each exactly once)
    for (int l = 2; l <= t.coarse_levels; ++l) {
      for (int row = 0; row < t.n; ++row)
        cmembers[gid(l, t.cmap[l - 1][row])].push_back(gid(1, t.cmap[0][row]));
      for (int j = 0; j < t.level_nodes[l]; ++j) {
        auto& m = cmembers[gid(l, j)];
        std::sort(m.begin(), m.end());
        m.erase(std::unique(m.begin(), m.end()), m.end());
      }
    }
    // banks
    const int bank_start = static_cast<int>(bank_sys.size());
    std::vector<double> stat = mas_static_blocks(t, Ks[ci]);
    size_t sb = 0;
    for (int l = 0; l <= t.coarse_levels; ++l) {
      for (int bk = 0; bk < t.level_banks[l]; ++bk, ++sb) {
        bank_sys.push_back(ci);
        bank_level.push_back(l);
        for (int s = 0; s < kBankSlots; ++s) {
          const int id = bk * kBankSlots + s;
          int m = -1;
          if (l == 0) {
            // padded slot -> row
            const auto& po = cloths_[ci].part.offsets;
            const int r0 = po[bk], r1 = po[bk + 1];
            m = (r0 + s < r1) ? vo + r0 + s : -1;
          } else if (id < t.level_nodes[l]) {
            m = gid(l, id);
          }
          bank_member.push_back(m);
        }
        for (int e = 0; e < 1024; ++e)
          bank_static.push_back(static_cast<float>(stat[sb * 1024 + e]));
      }
    }
    // Persistent-PCG view of this system: slots, system-local coarse ids
    // (level 1 first, then deeper levels), coarse banks and, per coarse
    // node of level >= 2, the level-1 nodes whose residual sums make up its
    // own.
    {
      const auto& po = cloths_[ci].part.offsets;
      const int nb0 = t.level_banks[0];
      PcgSysDev sd;
      sd.slot0 = static_cast<int>(ps_slot_row.size());
      sd.nslot = kBankSlots * nb0;
      sd.bank0 = bank_start;
      sd.cb0 = static_cast<int>(ps_cb_bank.size());
      sd.cn0 = static_cast<int>(ps_cn_ptr.size()) - 1;
      std::vector<int> lvl_off(t.coarse_levels + 2, 0);
      for (int l = 1; l <= t.coarse_levels; ++l)
        lvl_off[l + 1] = lvl_off[l] + t.level_nodes[l];
      sd.ncn = lvl_off[t.coarse_levels + 1];
      sd.n1 = t.coarse_levels >= 1 ? t.level_nodes[1] : 0;
      for (int bk = 0; bk < nb0; ++bk)
        for (int s = 0; s < kBankSlots; ++s) {
          const int lr = po[bk] + s;
          const bool real = lr < po[bk + 1];
          const int slot = static_cast<int>(ps_slot_row.size());
          ps_slot_row.push_back(real ? vo + lr : -1);
          if (real) ps_slot_of_row[vo + lr] = slot;
          for (int l = 1; l <= 4; ++l)
            ps_slot_cn.push_back(real && l <= t.coarse_levels
                                     ? lvl_off[l] + t.cmap[l - 1][lr]
                                     : -1);
        }
      int bidx = bank_start + nb0;
      for (int l = 1; l <= t.coarse_levels; ++l)
        for (int bk = 0; bk < t.level_banks[l]; ++bk, ++bidx) {
          ps_cb_bank.push_back(bidx);
          for (int s = 0; s < kBankSlots; ++s) {
            const int id = bk * kBankSlots + s;
            ps_cb_member.push_back(id < t.level_nodes[l] ? lvl_off[l] + id : -1);
          }
        }
      sd.ncb = static_cast<int>(ps_cb_bank.size()) - sd.cb0;
      // The level-1 nodes that make up every coarse node (a level-1 node is
      // its own list).
      std::vector<std::vector<int>> desc(sd.ncn);
      for (int m = 0; m < sd.n1; ++m) desc[m].push_back(m);
      for (int lr = 0; lr < t.n; ++lr)
        for (int l = 2; l <= t.coarse_levels; ++l)
          desc[lvl_off[l] + t.cmap[l - 1][lr]].push_back(lvl_off[1] +
                                                          t.cmap[0][lr]);
      for (int l = 0; l < 5; ++l)
        sd.lvl[l] = l + 1 <= t.coarse_levels + 1 ? lvl_off[l + 1] : sd.ncn;
      for (auto& dsc : desc) {
        std::sort(dsc.begin(), dsc.end());
        dsc.erase(std::unique(dsc.begin(), dsc.end()), dsc.end());
        ps_cn_list.insert(ps_cn_list.end(), dsc.begin(), dsc.end());
        ps_cn_ptr.push_back(static_cast<int>(ps_cn_list.size()));
      }
      const int per_wg = kPcgThreads / kBankSlots;
      sd.nwg = std::max(1, (nb0 + per_wg - 1) / per_wg);
      // Coarse banks are solved by workgroups of their own, one wavefront
      // per bank, in parallel with the slot workgroups' level-0 work; they
      // fill whole workgroups, as every workgroup of a system takes part in
      // its exchanges and an exchange slows with their number.
      sd.ncw = (sd.ncb + kPcgWaves - 1) / kPcgWaves;
      sd.wg0 = ps_wg;
      ps_wg += sd.nwg + sd.ncw;
      ps_sys.push_back(sd);
    }
  }
  // Slot-ordered sliced ELL of the off-diagonal operator (64-slot slices).
  std::vector<int> ps_sell_off, ps_sell_width, ps_sell_col;
  std::vector<float> ps_sell_val;
  {
    const int nslot = static_cast<int>(ps_slot_row.size());
    std::vector<int> row_sys(nv, 0);
    for (int ci = 0; ci < S.nsys; ++ci)
      for (int r = S.vert_off[ci]; r < S.vert_off[ci + 1]; ++r) row_sys[r] = ci;
    auto offdiag = [&](int row) {
      const CsrMatrix& K = Ks[row_sys[row]];
      const int lr = row - S.vert_off[row_sys[row]];
      int cnt = 0;
      for (int a = K.row_ptr[lr]; a < K.row_ptr[lr + 1]; ++a) cnt += K.col[a] != lr;
      return cnt;
    };
    for (int s0 = 0; s0 < nslot; s0 += kSlice) {
      int width = 0;
      for (int s = s0; s < std::min(nslot, s0 + kSlice); ++s)
        if (ps_slot_row[s] >= 0) width = std::max(width, offdiag(ps_slot_row[s]));
      // Whole chunks of neighbours: the kernel gathers them 8 at a time.
      width = (width + kPcgNbrChunk - 1) / kPcgNbrChunk * kPcgNbrChunk;
      ps_sell_off.push_back(static_cast<int>(ps_sell_col.size()));
      ps_sell_width.push_back(width);
      const size_t base = ps_sell_col.size();
      ps_sell_col.resize(base + static_cast<size_t>(width) * kSlice);
      ps_sell_val.resize(base + static_cast<size_t>(width) * kSlice);
      for (int lane = 0; lane < kSlice; ++lane) {
        const int s = s0 + lane;
        const int row = s < nslot ? ps_slot_row[s] : -1;
        int kk = 0;
        if (row >= 0) {
          const int ci = row_sys[row];
          const CsrMatrix& K = Ks[ci];
          const int lr = row - S.vert_off[ci];
          for (int a = K.row_ptr[lr]; a < K.row_ptr[lr + 1]; ++a) {
            if (K.col[a] == lr) continue;
            const size_t e = base + static_cast<size_t>(kk) * kSlice + lane;
            ps_sell_col[e] = ps_slot_of_row[S.vert_off[ci] + K.col[a]];
            ps_sell_val[e] = static_cast<float>(K.val[a]);
            ++kk;
          }
        }
        for (; kk < width; ++kk) {
          const size_t e = base + static_cast<size_t>(kk) * kSlice + lane;
          ps_sell_col[e] = s < nslot ? s : 0;
          ps_sell_val[e] = 0.0f;
        }
      }
    }
  }
  // Workgroup of every slot (the kernel's bank split), the other
  // workgroups' slots each one reads (its halo) and every neighbour as an
  // index into the workgroup's slots [0, kPcgThreads) then its halo.
  std::vector<int> ps_halo_ptr{0}, ps_halo_slot, ps_sell_lcol(ps_sell_col.size(), 0);
  std::vector<int> ps_wg_n1_ptr{0}, ps_wg_n1, ps_wg_c_ptr{0}, ps_wg_c;
  // Per coarse bank member: the positions [begin, end) of its level-1 nodes
  // in its coarse workgroup's list.
  std::vector<int> ps_cb_rng(2 * ps_cb_member.size() + 2, 0);
  int need_max = 0;
  int nhalo_max = 0, nwg_max = 0, nc_max = 0;
  bool xch_fits = true;
  // Coarse nodes of every slot's levels as positions in its owner
  // workgroup's list, and of every halo entry in the reading workgroup's.
  std::vector<int> ps_slot_cnl(ps_slot_cn.size(), -1), ps_halo_cnl;
  {
    const int nslot = static_cast<int>(ps_slot_row.size());
    std::vector<int> slot_wg(nslot, -1), slot_local(nslot, 0);
    for (const auto& sd : ps_sys) {
      nwg_max = std::max(nwg_max, sd.nwg + sd.ncw);
      const int nb0 = sd.nslot / kBankSlots;
      for (int j = 0; j < sd.nwg; ++j) {
        const int lo = sd.slot0 + kBankSlots * ((j * nb0) / sd.nwg);
        const int hi = sd.slot0 + kBankSlots * (((j + 1) * nb0) / sd.nwg);
        for (int sl = lo; sl < hi; ++sl) {
          slot_wg[sl] = sd.wg0 + j;
          slot_local[sl] = sl - lo;
        }
      }
    }
    std::vector<std::vector<int>> halo(ps_wg);
    auto for_entries = [&](auto&& fn) {
      for (size_t si = 0; si < ps_sell_off.size(); ++si) {
        const int s0 = static_cast<int>(si) * kSlice;
        for (int lane = 0; lane < kSlice && s0 + lane < nslot; ++lane)
          for (int k = 0; k < ps_sell_width[si]; ++k)
            fn(s0 + lane, static_cast<size_t>(ps_sell_off[si]) +
                              static_cast<size_t>(k) * kSlice + lane);
      }
    };
    for_entries([&](int sl, size_t e) {
      const int c = ps_sell_col[e];
      if (slot_wg[c] != slot_wg[sl]) halo[slot_wg[sl]].push_back(c);
    });
    for (auto& hl : halo) {
      std::sort(hl.begin(), hl.end());
      hl.erase(std::unique(hl.begin(), hl.end()), hl.end());
      nhalo_max = std::max(nhalo_max, static_cast<int>(hl.size()));
      ps_halo_slot.insert(ps_halo_slot.end(), hl.begin(), hl.end());
      ps_halo_ptr.push_back(static_cast<int>(ps_halo_slot.size()));
    }
    for_entries([&](int sl, size_t e) {
      const int c = ps_sell_col[e];
      const int w = slot_wg[sl];
      if (slot_wg[c] == w) {
        ps_sell_lcol[e] = slot_local[c];
      } else {
        const auto& hl = halo[w];
        ps_sell_lcol[e] = kPcgThreads + static_cast<int>(
            std::lower_bound(hl.begin(), hl.end(), c) - hl.begin());
      }
    });
    std::vector<std::vector<int>> wg_slots(ps_wg);
    for (int sl = 0; sl < nslot; ++sl)
      if (slot_wg[sl] >= 0) wg_slots[slot_wg[sl]].push_back(sl);
    ps_halo_cnl.assign(4 * ps_halo_slot.size(), -1);
    // Coarse bank cb of a system is solved by wavefront cb / ncw of its
    // coarse workgroup cb % ncw.  Per workgroup: the level-1 nodes whose sums
    // its banks read, member by member (a node read by several members is
    // listed for each), for coarse workgroups, and the coarse nodes whose
    // corrections its own and halo slots add, for slot workgroups.
    for (const auto& sd : ps_sys) {
      for (int j = 0; j < sd.nwg + sd.ncw; ++j) {
        const int w = sd.wg0 + j;
        const bool cwg = j >= sd.nwg;
        std::vector<int> n1s, cs;
        for (int cb = j - sd.nwg; cwg && cb < sd.ncb; cb += sd.ncw)
          for (int k = 0; k < kBankSlots; ++k) {
            const size_t mi = kBankSlots * static_cast<size_t>(sd.cb0 + cb) + k;
            const int m = ps_cb_member[mi];
            ps_cb_rng[2 * mi] = static_cast<int>(n1s.size());
            for (int e = m < 0 ? 0 : ps_cn_ptr[sd.cn0 + m];
                 m >= 0 && e < ps_cn_ptr[sd.cn0 + m + 1]; ++e)
              n1s.push_back(sd.cn0 + ps_cn_list[e]);
            ps_cb_rng[2 * mi + 1] = static_cast<int>(n1s.size());
          }
        auto add_slot = [&](int sl) {
          for (int l = 0; l < 4; ++l)
            if (ps_slot_cn[4 * static_cast<size_t>(sl) + l] >= 0)
              cs.push_back(ps_slot_cn[4 * static_cast<size_t>(sl) + l]);
        };
        for (int sl : wg_slots[w])
          if (ps_slot_row[sl] >= 0) add_slot(sl);
        for (int e = ps_halo_ptr[w]; e < ps_halo_ptr[w + 1]; ++e) add_slot(ps_halo_slot[e]);
        std::sort(cs.begin(), cs.end());
        cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
        nc_max = std::max(nc_max, static_cast<int>(cs.size()));
        auto pos = [&](int cn) {
          if (cn < 0) return -1;
          return static_cast<int>(std::lower_bound(cs.begin(), cs.end(), cn) - cs.begin());
        };
        for (int sl : wg_slots[w])
          if (ps_slot_row[sl] >= 0)
            for (int l = 0; l < 4; ++l)
              ps_slot_cnl[4 * static_cast<size_t>(sl) + l] =
                  pos(ps_slot_cn[4 * static_cast<size_t>(sl) + l]);
        for (int e = ps_halo_ptr[w]; e < ps_halo_ptr[w + 1]; ++e)
          for (int l = 0; l < 4; ++l)
            ps_halo_cnl[4 * static_cast<size_t>(e) + l] =
                pos(ps_slot_cn[4 * static_cast<size_t>(ps_halo_slot[e]) + l]);
        need_max = std::max(need_max, static_cast<int>(n1s.size()));
        ps_wg_n1.insert(ps_wg_n1.end(), n1s.begin(), n1s.end());
        ps_wg_n1_ptr.push_back(static_cast<int>(ps_wg_n1.size()));
        ps_wg_c.insert(ps_wg_c.end(), cs.begin(), cs.end());
        ps_wg_c_ptr.push_back(static_cast<int>(ps_wg_c.size()));
        // The halo slots a workgroup completes (the exchanges themselves
        // have no size limit: words beyond the ones kept in registers are
        // mapped when collected).
        const int nh = ps_halo_ptr[w + 1] - ps_halo_ptr[w];
        if (nh > kPcgThreads * kPcgHaloPerThread) xch_fits = false;
      }
    }
  }
  for (int g = 0; g < ncn; ++g)
    cnode_ptr[g + 1] = cnode_ptr[g] + static_cast<int>(cmembers[g].size());
  std::vector<int> cnode_list;
  cnode_list.reserve(cnode_ptr[ncn]);
  for (auto& m : cmembers) cnode_list.insert(cnode_list.end(), m.begin(), m.end());

  std::vector<int> vf_ptr(nv + 1, 0), vf_list;
  for (int v = 0; v < nv; ++v) {
    vf_ptr[v + 1] = vf_ptr[v] + static_cast<int>(vf[v].size());
    vf_list.insert(vf_list.end(), vf[v].begin(), vf[v].end());
  }

  // ---- upload
  S.x.upload(hx, st);
  S.v.upload(hv, st);
  for (DevBuf<float>* b : {&S.prev_x, &S.prev_v, &S.outer_x, &S.outer_v,
                           &S.inner_x, &S.pin_lambda, &S.dx})
    b->resize(3 * static_cast<size_t>(nv));
  S.elastic_rhs.resize(3 * static_cast<size_t>(nv));
  S.d.resize(nv);
  for (int i = 0; i < 2; ++i) {
    S.z[i].resize(6 * static_cast<size_t>(nf));
    S.u[i].resize(6 * static_cast<size_t>(nf));
  }
  S.F.upload(hF, st);
  S.vf_ptr.upload(vf_ptr, st);
  S.vf_list.upload(vf_list, st);
  S.pinned.upload(hpinned, st);
  S.fparam.upload(hfparam, st);
  S.fstiff.upload(hfstiff, st);
  S.mass.upload(hmass, st);
  S.damp.upload(hdamp, st);
  S.bend_rhs.upload(hbend, st);
  S.pin_target.upload(hpin_t, st);
  S.contact_base_w.upload(hcw, st);
  S.h_mass = hmass;
  S.h_pinned = hpinned;
  S.h_contact_w = hcw_d;
  S.admm_part.resize(kAdmmPartials * static_cast<size_t>(local_step_grid(nv)));

  S.nblk = static_cast<int>(blk_sys.size());
  S.nbank = static_cast<int>(bank_sys.size());
  S.h_bank_level = bank_level;
  S.nnz_off = 0;
  for (const auto& K : Ks)
    for (int r = 0; r < K.n; ++r)
      for (int a = K.row_ptr[r]; a < K.row_ptr[r + 1]; ++a) S.nnz_off += K.col[a] != r;
  S.blk_sys.upload(blk_sys, st);
  S.blk_row0.upload(blk_row0, st);
  S.blk_rows.upload(blk_rows, st);
  S.sys_blk0.upload(sys_blk0, st);
  S.h_blk0 = sys_blk0;
  S.slice_off.upload(slice_off, st);
  S.slice_width.upload(slice_width, st);
  if (ell_col.empty()) {
    ell_col.push_back(0);
    ell_val.push_back(0.0f);
  }
  S.ell_col.upload(ell_col, st);
  S.ell_val.upload(ell_val, st);
  S.bank_sys.upload(bank_sys, st);
  S.bank_level.upload(bank_level, st);
  S.bank_member.upload(bank_member, st);
  S.bank_static.upload(bank_static, st);
  S.bank_inv.resize(std::max<size_t>(1, bank_static.size()));
  S.bank_inv_h.resize(std::max<size_t>(1, bank_static.size()));
  S.bank_scale.resize(std::max<size_t>(1, bank_sys.size()));
  S.x0_copy.resize(3 * static_cast<size_t>(nv));
  S.r0d.resize(3 * static_cast<size_t>(nv));
  S.dxc.resize(3 * static_cast<size_t>(nv));
  S.sys_done.resize(static_cast<size_t>(std::max(1, S.nsys)));
  if (cnode_list.empty()) cnode_list.push_back(0);
  S.cnode_ptr.upload(cnode_ptr, st);
  S.cnode_list.upload(cnode_list, st);
  S.cnode_level.upload(cnode_level.empty() ? std::vector<int>{0} : cnode_level,
                       st);
  S.row_cnode.upload(row_cnode, st);
  S.r.resize(6 * static_cast<size_t>(nv));
  S.p.resize(3 * static_cast<size_t>(nv));
  S.q.resize(3 * static_cast<size_t>(nv));
  S.z_pcg.resize(3 * static_cast<size_t>(nv));
  S.zl0.resize(3 * static_cast<size_t>(nv));
  S.cr.resize(3 * static_cast<size_t>(std::max(1, ncn)));
  S.cz.resize(3 * static_cast<size_t>(std::max(1, ncn)));
  S.cd.resize(std::max(1, n1));
  S.part.resize(2 * static_cast<size_t>(std::max(1, S.nblk)));
  S.part2.resize(2 * static_cast<size_t>(std::max(1, S.nblk)));
  S.part3.resize(3 * static_cast<size_t>(std::max(1, S.nblk)));
  S.state.resize(2 * static_cast<size_t>(std::max(1, S.nsys)));
  // ---- persistent PCG
  {
    auto up = [&](DevBuf<int>& b, std::vector<int> v) {
      if (v.empty()) v.push_back(0);
      b.upload(v, st);
    };
    const int nslot = static_cast<int>(ps_slot_row.size());
    int max_ncb = 0, max_ncn = 0;
    for (const auto& sd : ps_sys) {
      max_ncb = std::max(max_ncb, sd.ncb);
      max_ncn = std::max(max_ncn, sd.ncn);
    }
    S.ps_sys.upload(ps_sys, st);
    up(S.ps_slot_row, ps_slot_row);
    up(S.ps_sell_off, ps_sell_off);
    up(S.ps_sell_width, ps_sell_width);
    up(S.ps_sell_lcol, ps_sell_lcol);
    up(S.ps_halo_ptr, ps_halo_ptr);
    up(S.ps_halo_slot, ps_halo_slot);
    up(S.ps_wg_n1_ptr, ps_wg_n1_ptr);
    up(S.ps_wg_n1, ps_wg_n1);
    up(S.ps_wg_c_ptr, ps_wg_c_ptr);
    up(S.ps_wg_c, ps_wg_c);
    if (ps_sell_val.empty()) ps_sell_val.push_back(0.0f);
    S.ps_sell_val.upload(ps_sell_val, st);
    up(S.ps_slot_cn, ps_slot_cn);
    up(S.ps_slot_cnl, ps_slot_cnl);
    if (ps_halo_cnl.empty()) ps_halo_cnl.push_back(-1);
    up(S.ps_halo_cnl, ps_halo_cnl);
    up(S.ps_cb_bank, ps_cb_bank);
    up(S.ps_cb_member, ps_cb_member);
    up(S.ps_cb_rng, ps_cb_rng);
    const int ncn_all = static_cast<int>(ps_cn_ptr.size()) - 1;
    S.ps_tz.resize(3 * static_cast<size_t>(std::max(1, nslot)));
    S.ps_tz.zero(st);
    S.ps_tdx.resize(3 * static_cast<size_t>(std::max(1, nslot)));
    S.ps_tdx.zero(st);
    S.ps_tn.resize(3 * static_cast<size_t>(std::max(1, ncn_all)));
    S.ps_tn.zero(st);
    S.ps_tp.resize(kPcgPartWords * static_cast<size_t>(std::max(1, ps_wg)));
    S.ps_tp.zero(st);
    S.ps_tc.resize(3 * static_cast<size_t>(std::max(1, ncn_all)));
    S.ps_tc.zero(st);
    S.ps_tcp.resize(2 * std::max<size_t>(1, ps_cb_bank.size()));
    S.ps_tcp.zero(st);
    S.ps_bar.resize(kPcgBarrierStride * static_cast<size_t>(std::max(1, S.nsys)));
    S.ps_bar.zero(st);
    S.ps_abort.resize(1);
    S.ps_abort.zero(st);
    S.ps_out.resize(std::max(1, S.nsys));
    S.admm_ctl.resize(1);
    PersistDev& a = S.pers;
    a.nsys = S.nsys;
    a.nslot_total = nslot;
    a.sys = S.ps_sys.get();
    a.slot_row = S.ps_slot_row.get();
    a.sell_off = S.ps_sell_off.get();
    a.sell_width = S.ps_sell_width.get();
    a.sell_lcol = S.ps_sell_lcol.get();
    a.halo_ptr = S.ps_halo_ptr.get();
    a.halo_slot = S.ps_halo_slot.get();
    a.wg_n1_ptr = S.ps_wg_n1_ptr.get();
    a.wg_n1 = S.ps_wg_n1.get();
    a.wg_c_ptr = S.ps_wg_c_ptr.get();
    a.wg_c = S.ps_wg_c.get();
    a.sell_val = S.ps_sell_val.get();
    a.slot_cn = S.ps_slot_cn.get();
    a.slot_cnl = S.ps_slot_cnl.get();
    a.halo_cnl = S.ps_halo_cnl.get();
    a.cb_bank = S.ps_cb_bank.get();
    a.cb_member = S.ps_cb_member.get();
    a.cb_rng = S.ps_cb_rng.get();
    a.bank_inv = S.bank_inv.get();
    a.bank_inv_h = S.bank_inv_h.get();
    a.bank_scale = S.bank_scale.get();
    a.sys_blk0 = S.sys_blk0.get();
    a.part2 = S.part2.get();
    a.d = S.d.get();
    a.r0 = S.r.get();
    a.x = S.inner_x.get();
    a.tz = S.ps_tz.get();
    a.tdx = S.ps_tdx.get();
    a.tn = S.ps_tn.get();
    a.tp = S.ps_tp.get();
    a.tc = S.ps_tc.get();
    a.tcp = S.ps_tcp.get();
    a.bar = S.ps_bar.get();
    a.abort_flag = S.ps_abort.get();
    a.out = S.ps_out.get();
    a.abs_tol = global_.linear_solver_abs_tol;
    a.max_iter = pcg_max_iter();
    a.need_max = need_max;
    a.ncn_max = max_ncn;
    a.nc_max = nc_max;
    a.ncb_max = max_ncb;
    a.nhalo_max = nhalo_max;
    a.nwg_max = nwg_max;
    a.grid = ps_wg;
    const char* mode = std::getenv("SILK_HIP_PCG");
    const bool want = !(mode && std::strcmp(mode, "launch") == 0);
    // Systems are launched in batches whose workgroups the device keeps
    // resident together: one launch when they all fit, else consecutive
    // launches of consecutive systems.  SILK_HIP_PCG_BATCH_WG caps a batch's
    // workgroups (forces several batches on a small scene).
    int cap_wg = 1 << 30;
    if (const char* e = std::getenv("SILK_HIP_PCG_BATCH_WG")) cap_wg = std::max(1, std::atoi(e));
    // The one-workgroup-per-CU variant first; when a system alone does not
    // fit the device that way, the variant that puts two on a CU.
    auto make_batches = [&](int dense) {
      PersistDev av = a;
      av.dense = dense;
      const size_t lds = pcg_persistent_lds(av);
      std::vector<std::pair<int, int>>& out = S.ps_variant[dense];
      out.clear();
      if (ps_wg <= 0) return false;
      for (int s0 = 0; s0 < S.nsys;) {
        const int w0 = ps_sys[s0].wg0;
        int s1 = s0, w1 = w0;
        while (s1 < S.nsys && ps_sys[s1].wg0 == w1) {
          const int we = w1 + ps_sys[s1].nwg + ps_sys[s1].ncw;
          if (we - w0 > cap_wg && s1 > s0) break;
          if (!pcg_persistent_fits(we - w0, lds, dense)) break;
          w1 = we;
          ++s1;
        }
        if (s1 == s0) {  // a system alone does not fit
          out.clear();
          return false;
        }
        out.emplace_back(w0, w1 - w0);
        s0 = s1;
      }
      return true;
    };
    const bool ok0 = make_batches(0);
    const char* dense_env = std::getenv("SILK_HIP_PCG_DENSE");
    const bool ok1 = !(dense_env && std::strcmp(dense_env, "0") == 0) && make_batches(1);
    if (!ok1) S.ps_variant[1].clear();
    a.dense = ok0 ? 0 : 1;
    const bool batches_ok = ok0 || ok1;
    S.ps_batches = S.ps_variant[a.dense];
    S.use_persistent = want && xch_fits && batches_ok;
  }
  S.coll = make_collision_world(*this, S);
  hip_check(hipStreamSynchronize(st), "scene upload");
  scene_ = std::move(sc);
  scene_dirty_ = false;
}

ClothDev Context::cloth_dev() {
  Scene& S = *scene_;
  ClothDev c;
  c.nv = S.nv;
  c.nf = S.nf;
  c.h = S.dt;
  for (int d = 0; d < 3; ++d) c.acc[d] = global_.acceleration[d];
  c.F = S.F.get();
  c.fparam = S.fparam.get();
  c.fstiff = S.fstiff.get();
  c.vf_ptr = S.vf_ptr.get();
  c.vf_list = S.vf_list.get();
  c.mass = S.mass.get();
  c.damp = S.damp.get();
  c.bend_rhs = S.bend_rhs.get();
  c.pinned = S.pinned.get();
  c.pin_target = S.pin_target.get();
  return c;
}

PcgDev Context::pcg_dev() {
  Scene& S = *scene_;
  PcgDev p;
  p.nsys = S.nsys;
  p.nv = S.nv;
  p.nblk = S.nblk;
  p.blk_sys = S.blk_sys.get();
  p.blk_row0 = S.blk_row0.get();
  p.blk_rows = S.blk_rows.get();
  p.sys_blk0 = S.sys_blk0.get();
  p.slice_off = S.slice_off.get();
  p.slice_width = S.slice_width.get();
  p.ell_col = S.ell_col.get();
  p.ell_val = S.ell_val.get();
  p.nbank = S.nbank;
  p.bank_sys = S.bank_sys.get();
  p.bank_level = S.bank_level.get();
  p.bank_member = S.bank_member.get();
  p.bank_static = S.bank_static.get();
  p.bank_inv = S.bank_inv.get();
  p.ncnode = S.ncnode;
  p.cnode_ptr = S.cnode_ptr.get();
  p.cnode_list = S.cnode_list.get();
  p.cnode_level = S.cnode_level.get();
  p.row_cnode = S.row_cnode.get();
  p.n1node = S.n1node;
  p.x = S.inner_x.get();
  p.dx = S.dx.get();
  p.r = S.r.get();
  p.p = S.p.get();
  p.q = S.q.get();
  p.z = S.z_pcg.get();
  p.zl0 = S.zl0.get();
  p.cr = S.cr.get();
  p.cz = S.cz.get();
  p.d = S.d.get();
  p.part = S.part.get();
  p.part2 = S.part2.get();
  p.state = S.state.get();
  p.abs_tol = global_.linear_solver_abs_tol;
  p.max_iter = pcg_max_iter();
  p.criterion = global_.pcg_criterion;
  p.mas_half = global_.mas_precision;
  p.bank_inv_h = S.bank_inv_h.get();
  p.bank_scale = S.bank_scale.get();
  return p;
}

// The persistent solve's arguments for the current global configuration.
static PersistDev persist_args(Scene& S, const silk_hip_global& g,
                               int max_iter) {
  PersistDev a = S.pers;
  a.abs_tol = g.linear_solver_abs_tol;
  a.max_iter = max_iter;
  a.criterion = g.pcg_criterion;
  a.mas_half = g.mas_precision;
  a.dx_out = g.true_residual_stats || g.pcg_criterion == SILK_HIP_PCG_TRUE
                 ? S.dx.get() : nullptr;
  return a;
}

void Context::launch_persistent(PersistDev a, float rel_tol) {
  for (const auto& [w0, n] : scene_->ps_batches) {
    a.wg_base = w0;
    a.grid = n;
    hip_check(launch_pcg_persistent(stream_, a, rel_tol), "persistent PCG launch");
  }
}

int Context::pcg_solve(float rel_tol, bool factor, const RhsArgs& ra,
                       silk_hip_step_stats* stats, bool restart) {
  Scene& S = *scene_;
  PcgDev p = pcg_dev();
  launch_pcg_residual(stream_, p, ra);
  if (factor) launch_mas_factor(stream_, p, S.cd.get());
  const size_t n3 = 3 * static_cast<size_t>(S.nv);
  if (global_.true_residual_stats || global_.pcg_criterion == SILK_HIP_PCG_TRUE)
    hip_check(hipMemcpyAsync(S.x0_copy.get(), S.inner_x.get(), n3 * sizeof(float),
                             hipMemcpyDeviceToDevice, stream_),
              "warm start copy");
  if (persistent_on()) {
    PersistDev a = persist_args(S, global_, pcg_max_iter());
    launch_persistent(a, rel_tol);
    S.h_state.resize(S.nsys);
    int abort_flag = 0;
    hip_check(hipMemcpyAsync(S.h_state.data(), S.ps_out.get(),
                             S.nsys * sizeof(PcgState), hipMemcpyDeviceToHost,
                             stream_),
              "pcg state");
    hip_check(hipMemcpyAsync(&abort_flag, S.ps_abort.get(), sizeof(int),
                             hipMemcpyDeviceToHost, stream_),
              "pcg abort flag");
    hip_check(hipStreamSynchronize(stream_), "pcg sync");
    if (abort_flag != 0) {
      // A barrier wait gave up; clear the barrier words for the next launch.
      S.ps_bar.zero(stream_);
      S.ps_abort.zero(stream_);
      hip_check(hipStreamSynchronize(stream_), "pcg barrier reset");
      throw HipError("persistent PCG: grid barrier timed out");
    }
  } else {
    pcg_iterate(rel_tol, ra);
  }
  int iters = 0;
  double rel = 0.0;
  for (const auto& s : S.h_state) {
    iters += s.iter;
    if (s.rr0 > 0) rel = std::max(rel, std::sqrt(s.rr / s.rr0));
    if (stats) {
      stats->pcg_max_solve_iterations = std::max(stats->pcg_max_solve_iterations, s.iter);
      stats->pcg_stalled_solves += s.status == 2 ? 1 : 0;
    }
  }
  return finish_pcg(iters, rel, ra, stats, restart);
}

void Context::verify_host(float rel_tol, const RhsArgs& ra, silk_hip_step_stats* stats) {
  Scene& S = *scene_;
  PcgDev p = pcg_dev();
  AdmmCtl* ctl = S.admm_ctl.get();
  for (int k = 0;; ++k) {
    const bool final = k == kPcgRestarts;
    launch_admm_init(stream_, ctl, rel_tol, rel_tol, rel_tol);
    launch_iterate_verify(stream_, p, ra, S.x0_copy.get(), S.dx.get(), S.part3.get(),
                          S.sys_done.get(), ctl, rel_tol, 1, final ? 1 : 0);
    hip_check(hipMemcpyAsync(S.h_admm_ctl, ctl, sizeof(AdmmCtl),
                             hipMemcpyDeviceToHost, stream_),
              "verify counters");
    hip_check(hipStreamSynchronize(stream_), "verify sync");
    const AdmmCtl hc = S.h_admm_ctl[0];
    if (hc.verify_done != 0 || final) {
      if (stats) {
        double it2;
        std::memcpy(&it2, &hc.max_true_bits, sizeof(double));
        stats->max_true_rel_residual =
            std::max(stats->max_true_rel_residual, static_cast<float>(std::sqrt(it2)));
        stats->pcg_unverified_solves += hc.pcg_unverified;
      }
      return;
    }
    // Every system runs again from its committed x (the multi-launch PCG
    // has no per-system skip); the next round tests the restart's iterate.
    pcg_solve(rel_tol, false, ra, stats, true);
  }
}

void Context::pcg_iterate(float rel_tol, const RhsArgs& ra) {
  (void)ra;
  Scene& S = *scene_;
  PcgDev p = pcg_dev();
  launch_pcg_start(stream_, p, rel_tol);
  int k = 0;
  int chunk = 8;
  const int max_iter = pcg_max_iter();
  for (;;) {
    launch_pcg_iters(stream_, p, k, chunk, rel_tol);
    k += chunk;
    S.h_state.resize(S.nsys);
    hip_check(hipMemcpyAsync(S.h_state.data(), S.state.get() + (k & 1) * S.nsys,
                             S.nsys * sizeof(PcgState), hipMemcpyDeviceToHost,
                             stream_),
              "pcg state");
    hip_check(hipStreamSynchronize(stream_), "pcg sync");
    bool all_done = true;
    for (const auto& s : S.h_state) all_done = all_done && s.done;
    if (all_done || k > max_iter + 2) break;
    chunk = 4;
  }
  launch_pcg_commit(stream_, p);
}

int Context::finish_pcg(int iters, double rel, const RhsArgs& ra,
                        silk_hip_step_stats* stats, bool restart) {
  Scene& S = *scene_;
  PcgDev p = pcg_dev();
  const bool verify = global_.pcg_criterion == SILK_HIP_PCG_TRUE;
  if (stats) {
    stats->pcg_iterations += iters;
    if (restart) stats->pcg_restarts += S.nsys;
    else stats->pcg_solves += 1;
    stats->last_pcg_rel_residual = static_cast<float>(rel);
    if (global_.true_residual_stats) {
      // x0 was copied before the solve; dx is the solve's correction.
      AdmmCtl* ctl = S.admm_ctl.get();
      hip_check(hipMemsetAsync(ctl, 0, sizeof(AdmmCtl), stream_), "ctl reset");
      RhsArgs rv = ra;
      rv.skip = nullptr;
      launch_solve_residuals(stream_, p, rv, S.x0_copy.get(), S.dx.get(),
                             S.inner_x.get(), S.part3.get(), ctl, verify ? 0 : 1);
      hip_check(hipMemcpyAsync(S.h_admm_ctl, ctl, sizeof(AdmmCtl),
                               hipMemcpyDeviceToHost, stream_),
                "solve residuals");
      hip_check(hipStreamSynchronize(stream_), "solve residuals");
      double it2, cm2;
      std::memcpy(&it2, &S.h_admm_ctl->max_true_bits, sizeof(double));
      std::memcpy(&cm2, &S.h_admm_ctl->max_commit_bits, sizeof(double));
      stats->max_true_rel_residual =
          std::max(stats->max_true_rel_residual, static_cast<float>(std::sqrt(it2)));
      stats->max_committed_rel_residual = std::max(
          stats->max_committed_rel_residual, static_cast<float>(std::sqrt(cm2)));
    }
  }
  return iters;
}

// ADMM loop driven from the device: every iteration's kernels are enqueued
// without waiting; k_admm_check runs the stopping test after each local step
// and later kernels become no-ops once it passes.  One synchronisation per
// solve reads the counters back.
bool Context::admm_solve_device(silk_hip_step_stats* stats) {
  Scene& S = *scene_;
  const auto& g = global_;
  // Learned warm start, or recording of training samples (from solves that
  // start at the previous iterate); both need each solve's correction.
  const bool learned = g.warm_start == SILK_HIP_WARM_LEARNED && has_warm_net();
  const bool extrap = g.warm_start == SILK_HIP_WARM_EXTRAPOLATE;
  const bool projected = learned || extrap;
  const bool record = recording() && g.warm_start == SILK_HIP_WARM_PREVIOUS;
  // The loop only enqueues kernels (no copies, fills or host waits), so it
  // can be captured.  Per-kernel timing records events between the launches,
  // and the warm-start paths advance host-side state per solve, so those run
  // without the graph.
  const int nit = g.max_inner_iteration;
  const bool graph = g.admm_graph == SILK_HIP_GRAPH_ON && !timing_admm() && !projected &&
                     !record && nit > 1;
  if (!graph) {
    admm_enqueue(projected, learned, record, 0, nit);
  } else {
    // The first iteration is enqueued eagerly, so the device works on it
    // (contact solve, local step) while the host captures the rest: the
    // capture would otherwise leave the device idle after every outer
    // iteration's synchronisation.
    admm_enqueue(projected, learned, record, 0, 1);
    hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed),
              "ADMM graph capture");
    try {
      admm_enqueue(projected, learned, record, 1, nit);
    } catch (...) {
      hipGraph_t gr = nullptr;
      (void)hipStreamEndCapture(stream_, &gr);
      if (gr != nullptr) (void)hipGraphDestroy(gr);
      (void)hipGetLastError();
      throw;
    }
    launch_admm_graph(stats);
  }
  AdmmCtl* ctl = S.admm_ctl.get();
  hip_check(hipMemcpyAsync(S.h_admm_ctl, ctl, sizeof(AdmmCtl),
                           hipMemcpyDeviceToHost, stream_),
            "admm ctl");
  hip_check(hipMemcpyAsync(S.h_admm_ctl + 1, S.ps_abort.get(), sizeof(int),
                           hipMemcpyDeviceToHost, stream_),
            "pcg abort flag");
  // Read by admm_device_finish after the next synchronisation of the stream
  // (the CCD's, when there is collision), so that the CCD is enqueued
  // behind the ADMM loop without a host round trip between them.
  admm_pending_ = true;
  return true;
}

void Context::launch_admm_graph(silk_hip_step_stats* stats) {
  Scene& S = *scene_;
  hipGraph_t gr = nullptr;
  hip_check(hipStreamEndCapture(stream_, &gr), "ADMM graph capture end");
  bool fresh = false;
  if (S.admm_exec != nullptr) {
    // Kernel arguments (contact counts, colour ranges, the z/u buffers of
    // the current parity) change between outer iterations; the update keeps
    // the executable graph when only arguments and launch shapes differ.
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult res = hipGraphExecUpdateSuccess;
    if (hipGraphExecUpdate(S.admm_exec, gr, &bad, &res) != hipSuccess ||
        res != hipGraphExecUpdateSuccess) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(S.admm_exec);
      S.admm_exec = nullptr;
    }
  }
  if (S.admm_exec == nullptr) {
    const hipError_t e = hipGraphInstantiate(&S.admm_exec, gr, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      S.admm_exec = nullptr;
      (void)hipGraphDestroy(gr);
      hip_check(e, "ADMM graph instantiate");
    }
    fresh = true;
  }
  (void)hipGraphDestroy(gr);
  hip_check(hipGraphLaunch(S.admm_exec, stream_), "ADMM graph launch");
  if (stats) {
    stats->graph_launches += 1;
    if (fresh) stats->graph_instantiations += 1;
  }
}

void Context::admm_enqueue(bool projected, bool learned, bool record, int it0, int it1) {
  Scene& S = *scene_;
  ClothDev c = cloth_dev();
  const size_t n3 = 3 * static_cast<size_t>(S.nv);
  const auto& g = global_;
  const bool first = it0 == 0;
  const bool last = it1 == g.max_inner_iteration;
  if (first) launch_fill_u32(stream_, S.pin_lambda.get(), static_cast<int64_t>(n3), 0u);
  ContactView cv = S.coll ? S.coll->contact_view() : ContactView{};
  AdmmCtl* ctl = S.admm_ctl.get();
  const int* skip = &ctl->done;
  if (first)
    launch_admm_init(stream_, ctl, g.initial_linear_rel_tol,
                     g.linear_solver_rel_tol_min, g.linear_solver_rel_tol_max);
  RhsArgs ra;
  ra.c = c;
  ra.prev_x = S.prev_x.get();
  ra.prev_v = S.prev_v.get();
  ra.elastic_rhs = S.elastic_rhs.get();
  ra.pin_lambda = S.pin_lambda.get();
  ra.cw = cv.w;
  ra.cp = cv.p;
  ra.ceta = cv.eta;
  ra.skip = skip;
  PcgDev p = pcg_dev();
  PersistDev pa = persist_args(S, g, pcg_max_iter());
  pa.ctl = ctl;
  const bool verify = g.pcg_criterion == SILK_HIP_PCG_TRUE;
  if (verify) {
    // The true-residual test runs inside each solve's launch: the residual
    // kernel keeps b - A x0 in FP64, the solve tests b - A (x0 + dx) when its
    // recurrence converges and replaces r by it when above the tolerance.
    ra.r0d = S.r0d.get();
    pa.r0d = S.r0d.get();
    pa.dxc = S.dxc.get();
    pa.max_replace = kPcgRestarts;
  }
  if (projected || record) {
    pa.dx_out = S.dx.get();
    const size_t ns = n3;
    if (S.wfeat.size() < ns * kMlpIn) {
      S.wfeat.resize(ns * kMlpIn);
      S.wsig.resize(ns);
      S.wdelta.resize(ns);
      S.w_xprev.resize(ns);
      S.w_rprev.resize(ns);
      S.w_r0dprev.resize(ns);
      S.w_part2.resize(2 * static_cast<size_t>(std::max(1, S.nblk)));
      S.w_gram.resize(30 * static_cast<size_t>(std::max(1, S.nblk)));
      S.w_coef.resize(6 * static_cast<size_t>(std::max(1, S.nsys)));
    }
    // The first solve of a segment has no previous correction.
    if (first) launch_fill_u32(stream_, S.dx.get(), static_cast<int64_t>(n3), 0u);
  }
  WarmFeatArgs wf;
  wf.r = S.r.get();
  wf.d = S.d.get();
  wf.dxp = S.dx.get();
  wf.v = S.outer_v.get();
  wf.cw = cv.w;
  wf.pinned = S.pinned.get();
  wf.h = S.dt;
  wf.skip = skip;
  AdmmCheckArgs ca;
  ca.grid = local_step_grid(S.nv);
  ca.n3 = static_cast<int>(n3);
  ca.primal_dof = 6 * S.nf + S.pinned_dof + (S.coll ? S.coll->active_dof() : 0);
  ca.abs_tol = g.admm_abs_tol;
  ca.rel_tol = g.admm_rel_tol;
  ca.lin_init = g.initial_linear_rel_tol;
  ca.lin_min = g.linear_solver_rel_tol_min;
  ca.lin_max = g.linear_solver_rel_tol_max;
  const bool keep_x0 = g.true_residual_stats != 0;
  for (int it = it0; it < it1; ++it) {
    // Without a projected start the residual kernel also keeps the solve's
    // start x for the true-residual test.
    const bool fold_copy = keep_x0 && it != 0 && !projected;
    if (it != 0) {
      RhsArgs r0 = ra;
      if (fold_copy) r0.x_copy = S.x0_copy.get();
      timed_begin(kTimeResidual);
      launch_pcg_residual(stream_, p, r0);
      timed_end();
      if (projected) {
        // Keep the previous iterate and its residual, project the start onto
        // span{network prediction, previous correction}, and let the judge
        // keep whichever of the two starts has the smaller FP64 residual.
        // (Doubles are copied as pairs of 32-bit words.)
        const int64_t n3w = static_cast<int64_t>(n3);
        launch_copy_unless(stream_, S.w_xprev.get(), S.inner_x.get(), n3w, skip);
        launch_copy_unless(stream_, S.w_rprev.get(), S.r.get(), n3w, skip);
        if (verify)
          launch_copy_unless(stream_, reinterpret_cast<float*>(S.w_r0dprev.get()),
                             reinterpret_cast<const float*>(S.r0d.get()), 2 * n3w, skip);
        launch_copy_unless(stream_, reinterpret_cast<float*>(S.w_part2.get()),
                           reinterpret_cast<const float*>(S.part2.get()),
                           4 * static_cast<int64_t>(S.nblk), skip);
        timed_begin(kTimeWarm);
        const float* v1 = nullptr;
        if (learned) {
          launch_warm_features(stream_, p, wf, S.wfeat.get(), S.wsig.get());
          launch_mlp_predict(stream_, warm_.dev(), S.wfeat.get(), S.wsig.get(),
                             static_cast<int64_t>(n3), nullptr, S.wdelta.get(), skip);
          v1 = S.wdelta.get();
        }
        launch_warm_project(stream_, p, v1, S.dx.get(), S.r.get(),
                            g.warm_projection == SILK_HIP_PROJECT_GALERKIN ? 0 : 1,
                            S.w_gram.get(), S.w_coef.get(), S.inner_x.get(), skip);
        timed_end();
        timed_begin(kTimeResidual);
        launch_pcg_residual(stream_, p, ra);
        timed_end();
        launch_warm_judge(stream_, p, S.w_part2.get(), S.w_xprev.get(), S.w_rprev.get(),
                          S.inner_x.get(), S.r.get(), &ctl->warm_kept, skip,
                          verify ? S.r0d.get() : nullptr,
                          verify ? S.w_r0dprev.get() : nullptr);
      } else if (record) {
        launch_warm_features(stream_, p, wf, S.wfeat.get(), S.wsig.get());
      }
      if (it == 1) launch_mas_factor(stream_, p, S.cd.get(), skip);
      if (keep_x0 && !fold_copy)
        launch_copy_unless(stream_, S.x0_copy.get(), S.inner_x.get(),
                           static_cast<int64_t>(n3), nullptr);
      timed_begin(kTimePcg);
      launch_persistent(pa, 0.0f);
      timed_end();
      // FP64 residuals of the iterate (folded here unless the solve's own
      // test folded it) and of the stored FP32 x.
      if (g.true_residual_stats)
        launch_solve_residuals(stream_, p, ra, S.x0_copy.get(), S.dx.get(),
                               S.inner_x.get(), S.part3.get(), ctl, verify ? 0 : 1);
      if (record)
        launch_warm_record(stream_, S.wfeat.get(), S.wsig.get(), S.dx.get(),
                           static_cast<int64_t>(n3), warm_.ds_rate, warm_.ds_salt++,
                           warm_.ds_feat.get(), warm_.ds_target.get(),
                           warm_.ds_count.get(), warm_.ds_cap, skip);
    }
    timed_begin(kTimeContact);
    if (S.coll) S.coll->contact_update(S.inner_x.get(), skip);
    timed_end();
    LocalStepArgs a;
    a.x = S.inner_x.get();
    a.z_in = S.z[S.aux_cur].get();
    a.u_in = S.u[S.aux_cur].get();
    a.z_out = S.z[S.aux_cur ^ 1].get();
    a.u_out = S.u[S.aux_cur ^ 1].get();
    a.elastic_rhs = S.elastic_rhs.get();
    a.pin_lambda = S.pin_lambda.get();
    a.contact_dual = cv.dual;
    a.contact_eta = cv.eta;
    a.contact_w = cv.w;
    a.contact_p = cv.p;
    a.partials = S.admm_part.get();
    a.skip = skip;
    timed_begin(kTimeLocal);
    launch_local_step(stream_, c, a);
    timed_end();
    S.aux_cur ^= 1;
    ca.it = it;
    launch_admm_check(stream_, S.admm_part.get(), ca, ctl);
  }
  if (last) launch_enforce(stream_, c, cv.w, cv.p, S.inner_x.get());
}

bool Context::admm_device_finish(silk_hip_step_stats* stats) {
  if (!admm_pending_) return true;
  admm_pending_ = false;
  Scene& S = *scene_;
  const auto& g = global_;
  const bool verify = g.pcg_criterion == SILK_HIP_PCG_TRUE;
  hip_check(hipStreamSynchronize(stream_), "admm sync");
  timed_collect();
  const AdmmCtl& hc = S.h_admm_ctl[0];
  last_admm_iters_ = hc.admm_iters;
  const int abort_flag = *reinterpret_cast<const int*>(S.h_admm_ctl + 1);
  if (abort_flag != 0) {
    S.ps_bar.zero(stream_);
    S.ps_abort.zero(stream_);
    hip_check(hipStreamSynchronize(stream_), "pcg barrier reset");
    throw HipError("persistent PCG: grid barrier timed out");
  }
  if (stats) {
    stats->admm_iterations += hc.admm_iters;
    stats->pcg_iterations += hc.pcg_iters;
    stats->pcg_solves += hc.pcg_solves;
    if (hc.pcg_solves > 0) stats->last_pcg_rel_residual = hc.last_rel;
    stats->pcg_max_solve_iterations =
        std::max(stats->pcg_max_solve_iterations, hc.pcg_max_solve);
    stats->pcg_stalled_solves += hc.pcg_stalled;
    stats->warm_start_kept += hc.warm_kept;
    stats->warm_start_solves += hc.warm_solves;
    stats->pcg_restarts += hc.pcg_restarts;
    stats->pcg_unverified_solves += hc.pcg_unverified;
    if (verify && !g.true_residual_stats) {
      double it2;
      std::memcpy(&it2, &hc.max_true_bits, sizeof(double));
      stats->max_true_rel_residual =
          std::max(stats->max_true_rel_residual, static_cast<float>(std::sqrt(it2)));
    }
    if (g.true_residual_stats) {
      double it2, cm2;
      std::memcpy(&it2, &hc.max_true_bits, sizeof(double));
      std::memcpy(&cm2, &hc.max_commit_bits, sizeof(double));
      stats->max_true_rel_residual =
          std::max(stats->max_true_rel_residual, static_cast<float>(std::sqrt(it2)));
      stats->max_committed_rel_residual = std::max(
          stats->max_committed_rel_residual, static_cast<float>(std::sqrt(cm2)));
    }
  }
  return hc.diverged == 0;
}

bool Context::admm_solve(silk_hip_step_stats* stats) {
  Scene& S = *scene_;
  if (persistent_on()) return admm_solve_device(stats);
  ClothDev c = cloth_dev();
  const size_t n3 = 3 * static_cast<size_t>(S.nv);
  launch_fill_u32(stream_, S.pin_lambda.get(), static_cast<int64_t>(n3), 0u);
  ContactView cv = S.coll ? S.coll->contact_view() : ContactView{};
  const int grid = local_step_grid(S.nv);
  double init_primal = 0.0, init_dual = 0.0, ratio = 1.0;
  int primal_dof = 6 * S.nf + S.pinned_dof;
  const auto& g = global_;
  RhsArgs ra;
  ra.c = c;
  ra.prev_x = S.prev_x.get();
  ra.prev_v = S.prev_v.get();
  ra.elastic_rhs = S.elastic_rhs.get();
  ra.pin_lambda = S.pin_lambda.get();
  ra.cw = cv.w;
  ra.cp = cv.p;
  ra.ceta = cv.eta;
  last_admm_iters_ = 0;
  for (int it = 0; it < g.max_inner_iteration; ++it) {
    if (it != 0) {
      const float rel = std::clamp(
          static_cast<float>(g.initial_linear_rel_tol * ratio),
          g.linear_solver_rel_tol_min, g.linear_solver_rel_tol_max);
      pcg_solve(rel, it == 1, ra, stats);
      if (g.pcg_criterion == SILK_HIP_PCG_TRUE) verify_host(rel, ra, stats);
    }
    if (S.coll) S.coll->contact_update(S.inner_x.get(), nullptr);
    LocalStepArgs a;
    a.x = S.inner_x.get();
    a.z_in = S.z[S.aux_cur].get();
    a.u_in = S.u[S.aux_cur].get();
    a.z_out = S.z[S.aux_cur ^ 1].get();
    a.u_out = S.u[S.aux_cur ^ 1].get();
    a.elastic_rhs = S.elastic_rhs.get();
    a.pin_lambda = S.pin_lambda.get();
    a.contact_dual = cv.dual;
    a.contact_eta = cv.eta;
    a.contact_w = cv.w;
    a.contact_p = cv.p;
    a.partials = S.admm_part.get();
    launch_local_step(stream_, c, a);
    S.aux_cur ^= 1;
    if (stats) stats->admm_iterations += 1;
    ++last_admm_iters_;
    if (it == 0) continue;

    download(S.h_admm_part, S.admm_part.get(),
             static_cast<size_t>(kAdmmPartials) * grid, stream_);
    hip_check(hipStreamSynchronize(stream_), "admm sync");
    double sum[kAdmmPartials] = {};
    for (int b = 0; b < grid; ++b)
      for (int j = 0; j < kAdmmPartials; ++j)
        sum[j] += S.h_admm_part[kAdmmPartials * b + j];
    if (sum[6] > 0.0) return false;  // non-finite state: diverged
    const int dof_p = primal_dof + (S.coll ? S.coll->active_dof() : 0);
    const double primal = std::sqrt(sum[0]);
    const double dual = std::sqrt(sum[3]);
    if (it == 1) {
      init_primal = primal;
      init_dual = dual;
    }
    const double eps_p = std::sqrt(static_cast<double>(dof_p)) * g.admm_abs_tol +
                         std::max(std::sqrt(sum[1]), std::sqrt(sum[2])) *
                             g.admm_rel_tol;
    const double eps_d =
        std::sqrt(static_cast<double>(3 * S.nv)) * g.admm_abs_tol +
        std::max(std::sqrt(sum[4]), std::sqrt(sum[5])) * g.admm_rel_tol;
    ratio = std::max(primal / std::max(init_primal, double(g.admm_abs_tol)),
                     dual / std::max(init_dual, double(g.admm_abs_tol)));
    if (primal <= eps_p && dual <= eps_d) break;
  }
  launch_enforce(stream_, c, cv.w, cv.p, S.inner_x.get());
  return true;
}

int Context::step(silk_hip_step_stats* stats) {
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (cloths_.empty()) {
    last_error_ = "nothing to solve";
    return SILK_HIP_ERR_UNKNOWN;
  }
  if (cpu_) return cpu_step(stats);
  if (scene_ && scene_->dt != global_.dt) invalidate_scene();
  if (scene_dirty_ || !scene_) build_scene();
  Scene& S = *scene_;
  ClothDev c = cloth_dev();
  const size_t n3 = 3 * static_cast<size_t>(S.nv);
  hip_check(hipEventRecord(ev_[0], stream_), "event");

  launch_step_init(stream_, c, S.x.get(), S.v.get(), S.prev_x.get(),
                   S.prev_v.get(), S.outer_x.get(), S.outer_v.get());
  // Aux variables restart every step (cloth_assembler.cu:137-142).
  launch_fill_u32(stream_, S.z[S.aux_cur].get(), static_cast<int64_t>(S.z[S.aux_cur].size()), 0u);
  launch_fill_u32(stream_, S.u[S.aux_cur].get(), static_cast<int64_t>(S.u[S.aux_cur].size()), 0u);

  float ms = 0.0f, dhat = 0.0f;
  if (S.coll) S.coll->begin_step(S.outer_x.get(), &ms, &dhat, stats);

  float remaining = 1.0f;
  float accepted_toi = 1.0f;
  toi_log_.clear();
  admm_log_.clear();
  int outer = 0;
  for (; outer < global_.max_outer_iteration; ++outer) {
    launch_predict(stream_, c, S.outer_x.get(), S.outer_v.get(),
                   S.inner_x.get());
    // The device ADMM loop is only enqueued here; its counters are read
    // after the CCD's synchronisation (a diverged solve is reported then,
    // and the CCD result on its positions discarded).
    bool ok = admm_solve(stats);
    float min_toi = remaining;
    if (S.coll && ok)
      min_toi = S.coll->min_toi(S.outer_x.get(), S.inner_x.get(),
                                1.0f - remaining, remaining, ms, stats);
    if (!admm_device_finish(stats) || !ok) {
      last_error_ = "solver diverged";
      return SILK_HIP_ERR_UNKNOWN;
    }
    if (S.coll) {
      toi_log_.push_back(min_toi);
      admm_log_.push_back(last_admm_iters_);
    }
    launch_velocity(stream_, static_cast<int>(n3), S.dt, S.outer_x.get(),
                    S.inner_x.get(), S.outer_v.get());
    if (min_toi >= remaining) {
      launch_copy_unless(stream_, S.outer_x.get(), S.inner_x.get(),
                         static_cast<int64_t>(n3), nullptr);
      break;
    }
    const float rollback = 0.8f * min_toi;
    accepted_toi = std::min(accepted_toi, 1.0f - remaining + min_toi);
    const float frac = rollback / remaining;
    launch_mix(stream_, static_cast<int>(n3), frac, S.inner_x.get(),
               S.outer_x.get(), S.outer_x.get());
    remaining -= rollback;
    if (S.coll) S.coll->rollback_contacts(S.outer_x.get(), 1.0f - remaining,
                                          dhat, stats);
  }
  // Velocity from the whole accepted displacement (main_loop.cu:281-296).
  launch_velocity(stream_, static_cast<int>(n3), S.dt, S.prev_x.get(),
                  S.outer_x.get(), S.v.get());
  launch_copy_unless(stream_, S.x.get(), S.outer_x.get(), static_cast<int64_t>(n3),
                     nullptr);
  if (S.coll) S.coll->end_step(stats);
  S.stepped = true;
  hip_check(hipEventRecord(ev_[1], stream_), "event");
  hip_check(hipStreamSynchronize(stream_), "step sync");
  float ms_step = 0.0f;
  hip_check(hipEventElapsedTime(&ms_step, ev_[0], ev_[1]), "elapsed");
  timing_ms_[0] += ms_step;
  timing_launch_[0] += 1;
  // Obstacle / pin static flags (main_loop.cu:298-306).
  for (auto& o : obstacles_) {
    if (o.has_pos) o.prev = o.curr;
    if (o.is_static) o.is_static_twice = true;
    else o.is_static = true;
  }
  if (stats) {
    stats->outer_iterations = std::min(outer + 1, global_.max_outer_iteration);
    stats->min_toi = accepted_toi;
  }
  return SILK_HIP_OK;
}

int Context::debug_pcg(const float* d, const float* b, float rel_tol, float* x,
                       float* launch_ms) {
  if (scene_ && scene_->dt != global_.dt) invalidate_scene();
  if (scene_dirty_ || !scene_) build_scene();
  Scene& S = *scene_;
  const size_t nv = static_cast<size_t>(S.nv), n3 = 3 * nv;
  PcgDev p = pcg_dev();
  // r0 = b (x0 = 0), and the per-block |r0|^2, |b|^2 partials the
  // right-hand-side criterion reads.
  std::vector<double> part2(2 * static_cast<size_t>(std::max(1, S.nblk)), 0.0);
  {
    std::vector<int> row0(S.nblk), rows(S.nblk);
    hip_check(hipMemcpyAsync(row0.data(), S.blk_row0.get(), S.nblk * sizeof(int),
                             hipMemcpyDeviceToHost, stream_),
              "blocks");
    hip_check(hipMemcpyAsync(rows.data(), S.blk_rows.get(), S.nblk * sizeof(int),
                             hipMemcpyDeviceToHost, stream_),
              "blocks");
    hip_check(hipStreamSynchronize(stream_), "blocks");
    for (int blk = 0; blk < S.nblk; ++blk) {
      double s = 0.0;
      for (int r = row0[blk]; r < row0[blk] + rows[blk]; ++r)
        for (int k = 0; k < 3; ++k) s += static_cast<double>(b[3 * r + k]) * b[3 * r + k];
      part2[2 * blk] = s;
      part2[2 * blk + 1] = s;
    }
  }
  hip_check(hipMemcpyAsync(S.d.get(), d, nv * sizeof(float), hipMemcpyHostToDevice,
                           stream_),
            "debug d");
  hip_check(hipMemcpyAsync(S.r.get(), b, n3 * sizeof(float), hipMemcpyHostToDevice,
                           stream_),
            "debug b");
  hip_check(hipMemcpyAsync(S.part2.get(), part2.data(), part2.size() * sizeof(double),
                           hipMemcpyHostToDevice, stream_),
            "debug partials");
  hip_check(hipMemsetAsync(S.inner_x.get(), 0, n3 * sizeof(float), stream_), "x0");
  launch_mas_factor(stream_, p, S.cd.get());
  int iters = 0;
  if (persistent_on()) {
    PersistDev a = persist_args(S, global_, pcg_max_iter());
    a.dx_out = nullptr;
    if (launch_ms) hip_check(hipEventRecord(ev_[0], stream_), "event");
    launch_persistent(a, rel_tol);
    if (launch_ms) hip_check(hipEventRecord(ev_[1], stream_), "event");
    S.h_state.resize(S.nsys);
    hip_check(hipMemcpyAsync(S.h_state.data(), S.ps_out.get(), S.nsys * sizeof(PcgState),
                             hipMemcpyDeviceToHost, stream_),
              "pcg state");
  } else {
    pcg_iterate(rel_tol, RhsArgs{});  // commits x += dx
  }
  hip_check(hipMemcpyAsync(x, S.inner_x.get(), n3 * sizeof(float), hipMemcpyDeviceToHost,
                           stream_),
            "debug x");
  hip_check(hipStreamSynchronize(stream_), "debug pcg");
  if (launch_ms && persistent_on())
    hip_check(hipEventElapsedTime(launch_ms, ev_[0], ev_[1]), "elapsed");
  for (const auto& st : S.h_state) iters += st.iter;
  return iters;
}

int Context::autotune(int32_t* report) {
  for (int k = 0; k < SILK_HIP_AUTOTUNE_REPORT; ++k) report[k] = -1;
  if (scene_dirty_ || !scene_ || !scene_->stepped) {
    last_error_ = "autotune needs a solver step on the current scene";
    return SILK_HIP_ERR_INVALID_CONFIG;
  }
  Scene& S = *scene_;
  // Broadphase: the depth to which every traversal job is split into
  // independent sub-tree items.  The candidate lists are sorted before use,
  // so every depth gives the same results.
  if (S.coll) {
    int best = S.coll->split_depth();
    float best_ms = INFINITY;
    for (int depth = 2; depth <= 7; ++depth) {
      S.coll->set_split_depth(depth);
      const float ms = S.coll->time_broadphase(5) / 5.0f;
      report[2 + depth - 2] = static_cast<int32_t>(std::lround(1000.0f * ms));
      if (ms < best_ms) {
        best_ms = ms;
        best = depth;
      }
    }
    S.coll->set_split_depth(best);
    report[0] = best;
  }
  // PCG: one or two persistent workgroups per CU.  Both run the same
  // arithmetic in the same order; a solve of the scene's matrix from x = 0
  // to 1e-6 on a fixed right-hand side times each.
  if (persistent_on()) {
    const size_t nv = static_cast<size_t>(S.nv), n3 = 3 * nv;
    std::vector<float> d(nv), b(n3), x(n3);
    hip_check(hipMemcpyAsync(d.data(), S.d.get(), nv * sizeof(float),
                             hipMemcpyDeviceToHost, stream_),
              "autotune d");
    hip_check(hipStreamSynchronize(stream_), "autotune d");
    uint32_t h = 12345u;
    for (auto& v : b) {
      h = h * 1664525u + 1013904223u;
      v = static_cast<float>(h >> 8) * (2.0f / 16777216.0f) - 1.0f;
    }
    const int32_t crit = global_.pcg_criterion;
    global_.pcg_criterion = SILK_HIP_PCG_RHS;
    const int keep = S.pers.dense;
    float best_ms = INFINITY;
    int best = keep;
    std::vector<float> x_first;
    bool agree = true;
    for (int dense = 0; dense < 2; ++dense) {
      if (S.ps_variant[dense].empty()) continue;
      S.pers.dense = dense;
      S.ps_batches = S.ps_variant[dense];
      float ms = INFINITY;
      for (int rep = 0; rep < 4; ++rep) {
        float t = 0.0f;
        debug_pcg(d.data(), b.data(), 1e-6f, x.data(), &t);
        if (rep > 0) ms = std::min(ms, t);
      }
      // The variants differ only in residency: their solutions agree bit for bit.
      if (x_first.empty()) x_first = x;
      else agree = agree && std::memcmp(x_first.data(), x.data(), n3 * sizeof(float)) == 0;
      report[8 + dense] = static_cast<int32_t>(std::lround(1000.0f * ms));
      if (ms < best_ms) {
        best_ms = ms;
        best = dense;
      }
    }
    global_.pcg_criterion = crit;
    S.pers.dense = best;
    S.ps_batches = S.ps_variant[best];
    report[1] = best;
    if (!agree) {
      last_error_ = "the persistent PCG variants gave different solutions";
      return SILK_HIP_ERR_UNKNOWN;
    }
  }
  return SILK_HIP_OK;
}

void Context::update_pin_targets(const ClothEntity& c) {
  if (cpu_world_) {
    cpu_world_->set_pin_targets(static_cast<int>(&c - cloths_.data()), c.pin_curr.data());
    return;
  }
  if (!scene_) return;
  const size_t ci = static_cast<size_t>(&c - cloths_.data());
  const int off = scene_->vert_off[ci];
  std::vector<float> t(3 * static_cast<size_t>(c.mesh.nv()), 0.0f);
  for (size_t i = 0; i < c.pins.size(); ++i) {
    const int lv = c.part.perm[c.pins[i]];
    for (int d = 0; d < 3; ++d) t[3 * lv + d] = c.pin_curr[3 * i + d];
  }
  hip_check(hipMemcpyAsync(scene_->pin_target.get() + 3 * static_cast<size_t>(off),
                           t.data(), t.size() * sizeof(float),
                           hipMemcpyHostToDevice, stream_),
            "pin targets");
  hip_check(hipStreamSynchronize(stream_), "pin targets sync");
}

void Context::get_positions(const ClothEntity& c, float* out) const {
  const int nv = c.mesh.nv();
  if (cpu_world_) {
    const std::vector<float>& x = cpu_world_->x(static_cast<int>(&c - cloths_.data()));
    std::memcpy(out, x.data(), x.size() * sizeof(float));
    return;
  }
  if (scene_) {
    size_t ci = static_cast<size_t>(&c - cloths_.data());
    const int off = scene_->vert_off[ci];
    std::vector<float> hx(3 * static_cast<size_t>(nv));
    hip_check(hipMemcpyAsync(hx.data(), scene_->x.get() + 3 * off,
                             hx.size() * sizeof(float), hipMemcpyDeviceToHost,
                             stream_),
              "get_positions");
    hip_check(hipStreamSynchronize(stream_), "get_positions sync");
    for (int v = 0; v < nv; ++v)
      for (int d = 0; d < 3; ++d) out[3 * v + d] = hx[3 * c.part.perm[v] + d];
    return;
  }
  const std::vector<float>& src = c.has_state ? c.x : c.mesh.V;
  std::memcpy(out, src.data(), src.size() * sizeof(float));
}

}  // namespace silk::amd
