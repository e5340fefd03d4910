// Learned CG warm start: a per-vertex network over the 1-ring of the PD
// operator that predicts the correction the next x-update solve will make.
//
// No counterpart in the reference, which starts every CG solve from the
// previous ADMM iterate (mas_cg_solver.cu:193-311, admm_solver.cu:272-292).
// The prediction is not added to the start as it is: a per-vertex prediction
// carries high-frequency error that the stiff operator amplifies, so its
// residual is nearly always worse than the previous iterate's.  Instead the
// start is x0 + c1 * prediction + c2 * (previous solve's correction), with
// (c1, c2) per cloth and coordinate from a 2x2 projection of the x-update
// system (minimal FP64 residual norm, or Galerkin); a noisy direction gets
// a small coefficient.  Plain PCG stays in charge: the projected start is
// kept only when its FP64 residual is below that of the previous iterate,
// and the solve then runs to the same stopping test either way.
//
// One sample per (vertex, coordinate): the operator acts on the three
// coordinates alike (vertex-scalar matrix times I3), so the network is shared
// by them.  Sample s = 3 * row + c matches the layout of the state vectors.
//
// Features of sample (i, c), with r the FP64-assembled initial residual, d the
// diagonal, dxp the previous solve's correction, v the segment velocity and
// N(i) the operator's neighbours of i:
//   0 r_ic / d_i            1 mean_{j in N(i)} r_jc / d_j
//   2 dxp_ic                3 mean_{j in N(i)} dxp_jc
//   4 h v_ic                5 contact weight_i / d_i
//   6 pinned_i              7 sum_{j in N(i)} |K_ij| / d_i
//   8..15 zero
// Features 0-4 are divided by sigma = max_k |f_k| (k < 5) and the network
// predicts the correction divided by sigma, so the prediction scales with
// its inputs.
//
// Network: 16 -> 64 -> 64 -> 1, SiLU after the two hidden layers, BF16
// weights and activations into the MFMA, FP32 accumulation, FP32 biases.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "hip/kernels.hpp"

namespace silk::amd {

constexpr int kMlpIn = 16;
constexpr int kMlpHidden = 64;
constexpr int kMlpFeatures = 8;  // features in use (the rest are zero)

/// FP32 master parameters, flat: W1 [64][16] (hidden x input), b1 [64],
/// W2 [64][64] (out x in), b2 [64], W3 [64], b3 [1].
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpHidden * kMlpIn;
constexpr int kMlpW2 = kMlpB1 + kMlpHidden;
constexpr int kMlpB2 = kMlpW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpW3 = kMlpB2 + kMlpHidden;
constexpr int kMlpB3 = kMlpW3 + kMlpHidden;
constexpr int kMlpParams = kMlpB3 + 1;

/// Device view of the network: BF16 copies of the weights (made by
/// launch_mlp_convert) and the FP32 master parameters (biases are read from
/// there).
struct MlpDev {
  const unsigned short* w1 = nullptr;  // [64][16]
  const unsigned short* w2 = nullptr;  // [64][64]
  const unsigned short* w3 = nullptr;  // [64]
  const float* params = nullptr;       // kMlpParams
};

/// Samples of a training batch or of a recorded data set: 16 BF16 features
/// and the target (correction / sigma, clipped) per sample.
struct MlpSamples {
  const unsigned short* feat = nullptr;  // n x 16
  const float* target = nullptr;         // n
  int64_t n = 0;
};

/// Inputs of the feature kernel for one solve.
struct WarmFeatArgs {
  const float* r = nullptr;      // 3 nv: initial residual (FP32 copy)
  const float* d = nullptr;      // nv: diagonal
  const float* dxp = nullptr;    // 3 nv: previous solve's correction
  const float* v = nullptr;      // 3 nv: segment velocity
  const float* cw = nullptr;     // nv: active contact weight (may be null)
  const int* pinned = nullptr;   // nv
  float h = 0.0f;
  const int* skip = nullptr;     // no-op when *skip != 0
};

/// feat (3 nv x 16 BF16) and sigma (3 nv) for every sample.
void launch_warm_features(hipStream_t s, const PcgDev& p, const WarmFeatArgs& a,
                          unsigned short* feat, float* sigma);

/// delta = sigma * net(feat) for every sample, added to x when x is not
/// null.  No-op when *skip != 0.
void launch_mlp_predict(hipStream_t s, const MlpDev& net, const unsigned short* feat,
                        const float* sigma, int64_t n, float* x, float* delta,
                        const int* skip);

/// Projected start: per system and coordinate, x += c1 v1 + c2 v2 with
/// (c1, c2) from the 2x2 projection of the x-update system onto span{v1,
/// v2} (v1: the network's prediction, v2: the previous solve's correction;
/// either may be null).  mode 0: Galerkin (minimal A-norm error); mode 1:
/// minimal FP64 residual norm.  part: 30 doubles per row block; coef: 6
/// floats per system.  No-op when *skip != 0.
void launch_warm_project(hipStream_t s, const PcgDev& p, const float* v1, const float* v2,
                         const float* r, int mode, double* part, float* coef, float* x,
                         const int* skip);

/// Per system: keep the predicted start when the FP64 residual of x0 + delta
/// (part2 now) is below that of x0 (part2_prev, saved before the
/// prediction); otherwise restore x and r (and the FP64 residual r0d from
/// r0d_prev when both are given).  Counts the kept predictions in
/// counters[0] and the solves in counters[1].
void launch_warm_judge(hipStream_t s, const PcgDev& p, const double* part2_prev,
                       const float* x_prev, const float* r_prev, float* x,
                       float* r, int* counters, const int* skip,
                       double* r0d = nullptr, const double* r0d_prev = nullptr);

/// Appends samples to a data set (feat, clipped dx / sigma) with probability
/// `rate` per sample (hash of sample and `salt`), up to `cap` samples.
void launch_warm_record(hipStream_t s, const unsigned short* feat,
                        const float* sigma, const float* dx, int64_t n, float rate,
                        uint32_t salt, unsigned short* ds_feat, float* ds_target,
                        unsigned long long* ds_count, int64_t cap, const int* skip);

/// BF16 copies of the weights from the FP32 master parameters.
void launch_mlp_convert(hipStream_t s, const float* params, unsigned short* w1,
                        unsigned short* w2, unsigned short* w3);

/// Forward and backward over a batch: `batch` samples drawn from `data`
/// (sample i of step `step` is data index hash(i, step) % n; with
/// step < 0, samples 0 .. batch-1 in order).  Loss = mean (y - target)^2.
/// Per-wavefront gradient partials go to grad_part (kMlpParams each, at most
/// `waves` wavefronts) and loss partials to loss_part; y (if non-null)
/// receives the predictions of the samples in order.
void launch_mlp_train_pass(hipStream_t s, const MlpDev& net, const MlpSamples& data,
                           int64_t batch, int64_t step, float* grad_part,
                           double* loss_part, int waves, float* y);

/// Sums the gradient partials (fixed order), then with lr > 0 takes one Adam
/// step on params (m, v: moments; t: step number from 1); the summed
/// gradient is left in grad (kMlpParams) and the summed loss in *loss.
void launch_mlp_reduce_adam(hipStream_t s, const float* grad_part,
                            const double* loss_part, int waves, float* grad,
                            double* loss, float* params, float* m, float* v,
                            float lr, int t);

/// Wavefronts a training pass uses for `batch` samples.
int mlp_train_waves(int64_t batch);

}  // namespace silk::amd
