// PD / ADMM kernels of the cloth step (integrator, local projection,
// main-update assembly, constraint enforcement).
//
// Layout: positions and all per-dof vectors are float[3*nv] in solver (MAS
// partition) order; per-vertex scalars are float[nv]; faces are int[3*nf]
// with a per-face float4 (ia, ib, ic, sqrt(area)) that replaces the
// reference's 54-float Jacobian (cloth_assembly_l1_cache.cu:280-289).
#include "hip/device_common.hpp"
#include "hip/kernels.hpp"

namespace silk::amd {

using namespace dev;

namespace {

__global__ void k_step_init(ClothDev c, const float* __restrict__ x,
                            const float* __restrict__ v,
                            float* __restrict__ prev_x,
                            float* __restrict__ prev_v,
                            float* __restrict__ outer_x,
                            float* __restrict__ outer_v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.nv) return;
  const float dmp = c.damp[i];
  for (int d = 0; d < 3; ++d) {
    const float xi = x[3 * i + d];
    const float vi = dmp * v[3 * i + d];
    prev_x[3 * i + d] = xi;
    prev_v[3 * i + d] = vi;
    outer_x[3 * i + d] = xi;
    outer_v[3 * i + d] = vi;
  }
}

__global__ void k_predict(ClothDev c, const float* __restrict__ ox,
                          const float* __restrict__ ov,
                          float* __restrict__ inner) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.nv) return;
  const float h = c.h;
  for (int d = 0; d < 3; ++d) {
    float nx = ox[3 * i + d] + h * ov[3 * i + d];
    inner[3 * i + d] = nx + (h * h) * c.acc[d];
  }
}

__global__ void k_velocity(int n3, float h, const float* __restrict__ a,
                           const float* __restrict__ b,
                           float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  out[i] = (b[i] - a[i]) / h;
}

__global__ void k_mix(int n3, float w, const float* __restrict__ a,
                      const float* __restrict__ b, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  out[i] = w * a[i] + (1.0f - w) * b[i];
}

// Symmetric 2x2 eigen-decomposition based projection of a 3x2 matrix Y onto
// { U diag(s) V^T : s in [lo, hi]^2 }.  Closed form (no iterative SVD):
// G = Y^T Y = V diag(l) V^T, sigma = sqrt(l), proj = Y V diag(c(s)/s) V^T for
// well separated sigma, with an explicit orthonormal completion when Y is
// rank deficient.  Replaces svd32 (svd.cu:373-490) as used in
// cloth_admm_helper.cu:19-51.
__device__ void clamp_singular_values(const float y[6], float lo, float hi,
                                      float out[6]) {
  const f3 c0 = {y[0], y[1], y[2]};
  const f3 c1 = {y[3], y[4], y[5]};
  const float g00 = dot3(c0, c0), g01 = dot3(c0, c1), g11 = dot3(c1, c1);
  const float half_diff = 0.5f * (g00 - g11);
  const float disc = sqrtf(half_diff * half_diff + g01 * g01);
  const float mean = 0.5f * (g00 + g11);
  const float l1 = mean + disc;
  const float l2 = fmaxf(mean - disc, 0.0f);
  // Eigenvector of l1.
  float v1x, v1y;
  if (disc == 0.0f) {
    v1x = 1.0f;
    v1y = 0.0f;
  } else if (half_diff >= 0.0f) {
    // (l1 - g11, g01) has the larger first entry.
    v1x = half_diff + disc;
    v1y = g01;
  } else {
    v1x = g01;
    v1y = disc - half_diff;
  }
  const float inv_len = 1.0f / sqrtf(v1x * v1x + v1y * v1y);
  v1x *= inv_len;
  v1y *= inv_len;
  const float v2x = -v1y, v2y = v1x;
  (void)l1;
  (void)l2;
  // Singular values from |Y v| (accurate even when l2 suffers cancellation).
  f3 yv1 = v1x * c0 + v1y * c1;
  f3 yv2 = v2x * c0 + v2y * c1;
  const float s1 = sqrtf(dot3(yv1, yv1)), s2 = sqrtf(dot3(yv2, yv2));
  const float tiny = 1e-12f;
  f3 u1, u2;
  if (s1 > tiny) {
    u1 = (1.0f / s1) * yv1;
  } else {
    u1 = {1.0f, 0.0f, 0.0f};
  }
  if (s2 > 1e-6f * s1 && s2 > tiny) {
    u2 = (1.0f / s2) * yv2;
  } else {
    // Completion: the coordinate axis least aligned with u1 (lowest index on
    // ties), orthogonalised against u1.
    const float ax = fabsf(u1.x), ay = fabsf(u1.y), az = fabsf(u1.z);
    int axis = 0;
    float amin = ax;
    if (ay < amin) { axis = 1; amin = ay; }
    if (az < amin) { axis = 2; }
    const f3 e = {axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f,
                  axis == 2 ? 1.0f : 0.0f};
    f3 t = e - dot3(e, u1) * u1;
    u2 = (1.0f / sqrtf(dot3(t, t))) * t;
  }
  const float cs1 = fminf(fmaxf(s1, lo), hi);
  const float cs2 = fminf(fmaxf(s2, lo), hi);
  // proj = cs1 u1 v1^T + cs2 u2 v2^T
  f3 p0 = (cs1 * v1x) * u1 + (cs2 * v2x) * u2;
  f3 p1 = (cs1 * v1y) * u1 + (cs2 * v2y) * u2;
  out[0] = p0.x; out[1] = p0.y; out[2] = p0.z;
  out[3] = p1.x; out[4] = p1.y; out[5] = p1.z;
}

struct FaceUpdate {
  float sx[6];
  float z[6];
  float u[6];
};

// Per-face ADMM z/u update (cloth_admm_helper.cu:80-112,142-243).
__device__ FaceUpdate face_update(const ClothDev& c, const LocalStepArgs& a,
                                  int f) {
  const int i0 = c.F[3 * f], i1 = c.F[3 * f + 1], i2 = c.F[3 * f + 2];
  const f3 x0 = ld3(a.x, i0), x1 = ld3(a.x, i1), x2 = ld3(a.x, i2);
  const float ia = c.fparam[4 * f], ib = c.fparam[4 * f + 1],
              ic = c.fparam[4 * f + 2], w = c.fparam[4 * f + 3];
  const float k = c.fstiff[2 * f], rho = c.fstiff[2 * f + 1];
  const f3 e0 = x1 - x0, e1 = x2 - x0;
  const f3 F0 = ia * e0;
  const f3 F1 = ib * e0 + ic * e1;
  FaceUpdate o;
  o.sx[0] = F0.x; o.sx[1] = F0.y; o.sx[2] = F0.z;
  o.sx[3] = F1.x; o.sx[4] = F1.y; o.sx[5] = F1.z;
  float u[6];
  for (int j = 0; j < 6; ++j) u[j] = a.u_in[6 * f + j];
  float y[6];
  const float inv_pw = 1.0f / (rho * w);
  for (int j = 0; j < 6; ++j) y[j] = o.sx[j] + inv_pw * u[j];
  float proj[6];
  clamp_singular_values(y, 0.9f, 1.1f, proj);
  const float r = k / (k + rho);
  for (int j = 0; j < 6; ++j) o.z[j] = (1.0f - r) * y[j] + r * proj[j];
  const float wp = w * rho;
  for (int j = 0; j < 6; ++j) {
    float t = u[j] + (wp * o.sx[j] + (-wp) * o.z[j]);
    o.u[j] = fminf(fmaxf(t, -a.max_mul), a.max_mul);
  }
  return o;
}

__global__ void __launch_bounds__(kBlock)
    k_local_step(ClothDev c, LocalStepArgs a) {
  __shared__ double lds[17];
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (a.skip != nullptr && *a.skip != 0) {
    // The loop has stopped: carry the aux variables over to the buffers the
    // next iteration (or the next segment) reads.
    if (v < c.nv)
      for (int e = c.vf_ptr[v]; e < c.vf_ptr[v + 1]; ++e) {
        if ((c.vf_list[e] & 3) != 0) continue;
        const int f = c.vf_list[e] >> 2;
        for (int j = 0; j < 6; ++j) {
          a.z_out[6 * f + j] = a.z_in[6 * f + j];
          a.u_out[6 * f + j] = a.u_in[6 * f + j];
        }
      }
    return;
  }
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  if (v < c.nv) {
    double rhs[3] = {0.0, 0.0, 0.0};
    f3 dcur = {0, 0, 0}, dprev = {0, 0, 0};
    for (int e = c.vf_ptr[v]; e < c.vf_ptr[v + 1]; ++e) {
      const int f = c.vf_list[e] >> 2;
      const int corner = c.vf_list[e] & 3;
      const FaceUpdate o = face_update(c, a, f);
      const float ia = c.fparam[4 * f], ib = c.fparam[4 * f + 1],
                  ic = c.fparam[4 * f + 2], w = c.fparam[4 * f + 3];
      const float rho = c.fstiff[2 * f + 1];
      float g0, g1;
      if (corner == 0) {
        g0 = -ia;
        g1 = -(ib + ic);
      } else if (corner == 1) {
        g0 = ia;
        g1 = ib;
      } else {
        g0 = 0.0f;
        g1 = ic;
      }
      const f3 zn0 = {o.z[0], o.z[1], o.z[2]}, zn1 = {o.z[3], o.z[4], o.z[5]};
      const f3 zo0 = ld3(a.z_in + 6 * f, 0), zo1 = ld3(a.z_in + 6 * f, 1);
      const f3 un0 = {o.u[0], o.u[1], o.u[2]}, un1 = {o.u[3], o.u[4], o.u[5]};
      const float rw2 = rho * w * w;
      const f3 szn = g0 * zn0 + g1 * zn1;
      const f3 szo = g0 * zo0 + g1 * zo1;
      const f3 sun = g0 * un0 + g1 * un1;
      dcur = dcur + rw2 * szn;
      dprev = dprev + rw2 * szo;
      rhs[0] += static_cast<double>(rw2) * szn.x - static_cast<double>(w) * sun.x;
      rhs[1] += static_cast<double>(rw2) * szn.y - static_cast<double>(w) * sun.y;
      rhs[2] += static_cast<double>(rw2) * szn.z - static_cast<double>(w) * sun.z;
      if (corner == 0) {
        // Face owner: write the new aux variables and the primal terms.
        for (int j = 0; j < 6; ++j) {
          a.z_out[6 * f + j] = o.z[j];
          a.u_out[6 * f + j] = o.u[j];
          const float pr = w * (o.sx[j] - o.z[j]);
          const float px = w * o.sx[j];
          const float pz = w * o.z[j];
          s[0] += static_cast<double>(pr * pr);
          s[1] += static_cast<double>(px * px);
          s[2] += static_cast<double>(pz * pz);
        }
      }
    }
    a.elastic_rhs[3 * v] = rhs[0];
    a.elastic_rhs[3 * v + 1] = rhs[1];
    a.elastic_rhs[3 * v + 2] = rhs[2];
    const f3 xv = ld3(a.x, v);
    // Pins (equality_constraints.cu:90-154).
    if (c.pinned[v]) {
      const f3 t = ld3(c.pin_target, v);
      const float xs[3] = {xv.x, xv.y, xv.z};
      const float ts[3] = {t.x, t.y, t.z};
      for (int d = 0; d < 3; ++d) {
        float lam = a.pin_lambda[3 * v + d] + a.pin_rho * (xs[d] - ts[d]);
        a.pin_lambda[3 * v + d] = fminf(fmaxf(lam, -a.max_mul), a.max_mul);
        const float diff = xs[d] - ts[d];
        s[0] += static_cast<double>(diff * diff);
        s[1] += static_cast<double>(xs[d] * xs[d]);
        s[2] += static_cast<double>(ts[d] * ts[d]);
      }
    }
    f3 dual = dcur - dprev;
    // Contact terms (contact_constraints.cu:316-364).
    if (a.contact_w != nullptr && a.contact_w[v] > 0.0f) {
      const f3 pv = ld3(a.contact_p, v);
      const f3 diff = xv - pv;
      s[0] += static_cast<double>(dot3(diff, diff));
      dual = dual + ld3(a.contact_dual, v);
      const f3 eta = ld3(a.contact_eta, v);
      dcur = dcur + eta;
      dprev = dprev + eta;
    }
    s[3] = static_cast<double>(dual.x) * dual.x +
           static_cast<double>(dual.y) * dual.y +
           static_cast<double>(dual.z) * dual.z;
    s[4] = static_cast<double>(dcur.x) * dcur.x +
           static_cast<double>(dcur.y) * dcur.y +
           static_cast<double>(dcur.z) * dcur.z;
    s[5] = static_cast<double>(dprev.x) * dprev.x +
           static_cast<double>(dprev.y) * dprev.y +
           static_cast<double>(dprev.z) * dprev.z;
    s[6] = (isfinite(xv.x) && isfinite(xv.y) && isfinite(xv.z)) ? 0.0 : 1.0;
  }
  for (int j = 0; j < 7; ++j) {
    double t = block_sum(s[j], lds);
    if (threadIdx.x == 0) a.partials[kAdmmPartials * blockIdx.x + j] = t;
  }
  if (threadIdx.x == 0) a.partials[kAdmmPartials * blockIdx.x + 7] = 0.0;
}

__global__ void k_enforce(ClothDev c, const float* __restrict__ cw,
                          const float* __restrict__ cp, float* __restrict__ x) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= c.nv) return;
  if (c.pinned[v]) st3(x, v, ld3(c.pin_target, v));
  if (cw != nullptr && cw[v] > 0.0f) st3(x, v, ld3(cp, v));
}

inline int grid_for(int n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

void launch_step_init(hipStream_t s, const ClothDev& c, const float* x,
                      const float* v, float* prev_x, float* prev_v,
                      float* outer_x, float* outer_v) {
  if (c.nv == 0) return;
  k_step_init<<<grid_for(c.nv), kBlock, 0, s>>>(c, x, v, prev_x, prev_v,
                                                outer_x, outer_v);
}

void launch_predict(hipStream_t s, const ClothDev& c, const float* outer_x,
                    const float* outer_v, float* inner) {
  if (c.nv == 0) return;
  k_predict<<<grid_for(c.nv), kBlock, 0, s>>>(c, outer_x, outer_v, inner);
}

void launch_velocity(hipStream_t s, int n3, float h, const float* a_start,
                     const float* b_end, float* out) {
  if (n3 == 0) return;
  k_velocity<<<grid_for(n3), kBlock, 0, s>>>(n3, h, a_start, b_end, out);
}

void launch_mix(hipStream_t s, int n3, float w, const float* a, const float* b,
                float* out) {
  if (n3 == 0) return;
  k_mix<<<grid_for(n3), kBlock, 0, s>>>(n3, w, a, b, out);
}

// Stopping test of admm_solver.cu:296-330 on the device; one workgroup.
__global__ void __launch_bounds__(kBlock)
    k_admm_check(const double* __restrict__ part, AdmmCheckArgs a, AdmmCtl* ctl) {
  __shared__ double lds[17];
  __shared__ double tot[kAdmmPartials];
  if (ctl->done) return;
  for (int j = 0; j < 7; ++j) {
    double v = 0.0;
    for (int b = threadIdx.x; b < a.grid; b += blockDim.x)
      v += part[kAdmmPartials * b + j];
    const double s = block_sum(v, lds);
    if (threadIdx.x == 0) tot[j] = s;
  }
  if (threadIdx.x != 0) return;
  AdmmCtl c = *ctl;
  c.admm_iters += 1;
  if (a.it == 0) {
    *ctl = c;
    return;
  }
  if (tot[6] > 0.0) {
    c.diverged = 1;
    c.done = 1;
    *ctl = c;
    return;
  }
  const double primal = sqrt(tot[0]);
  const double dual = sqrt(tot[3]);
  if (a.it == 1) {
    c.init_primal = primal;
    c.init_dual = dual;
  }
  const double at = a.abs_tol, rt = a.rel_tol;
  const double eps_p = sqrt(static_cast<double>(a.primal_dof)) * at +
                       fmax(sqrt(tot[1]), sqrt(tot[2])) * rt;
  const double eps_d = sqrt(static_cast<double>(a.n3)) * at +
                       fmax(sqrt(tot[4]), sqrt(tot[5])) * rt;
  c.ratio = fmax(primal / fmax(c.init_primal, at), dual / fmax(c.init_dual, at));
  const float rel = static_cast<float>(static_cast<double>(a.lin_init) * c.ratio);
  c.rel_tol = fminf(fmaxf(rel, a.lin_min), a.lin_max);
  if (primal <= eps_p && dual <= eps_d) c.done = 1;
  *ctl = c;
}

__global__ void k_admm_init(AdmmCtl* ctl, float lin_init, float lin_min,
                            float lin_max) {
  AdmmCtl c;
  c.rel_tol = fminf(fmaxf(lin_init, lin_min), lin_max);
  *ctl = c;
}

int local_step_grid(int nv) { return grid_for(nv); }

void launch_admm_init(hipStream_t s, AdmmCtl* ctl, float lin_init,
                      float lin_min, float lin_max) {
  k_admm_init<<<1, 1, 0, s>>>(ctl, lin_init, lin_min, lin_max);
}

void launch_admm_check(hipStream_t s, const double* partials,
                       const AdmmCheckArgs& a, AdmmCtl* ctl) {
  k_admm_check<<<1, kBlock, 0, s>>>(partials, a, ctl);
}

void launch_local_step(hipStream_t s, const ClothDev& c,
                       const LocalStepArgs& a) {
  if (c.nv == 0) return;
  k_local_step<<<grid_for(c.nv), kBlock, 0, s>>>(c, a);
}

void launch_enforce(hipStream_t s, const ClothDev& c, const float* contact_w,
                    const float* contact_p, float* x) {
  if (c.nv == 0) return;
  k_enforce<<<grid_for(c.nv), kBlock, 0, s>>>(c, contact_w, contact_p, x);
}

}  // namespace silk::amd
