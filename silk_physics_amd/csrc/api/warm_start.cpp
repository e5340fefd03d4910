// Learned CG warm start: network state, training loop and sample recording
// of a context (the kernels are in hip/warm_start.hip).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "api/context.hpp"

namespace silk::amd {

namespace {

uint64_t next_u64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Standard normal by Box-Muller from two uniforms in (0, 1].
double next_normal(uint64_t& s) {
  const double u1 = (static_cast<double>(next_u64(s) >> 11) + 1.0) * 0x1.0p-53;
  const double u2 = static_cast<double>(next_u64(s) >> 11) * 0x1.0p-53;
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// He-scaled initialisation of the two hidden layers and a small output layer
// (the untrained network predicts nearly no correction).
std::vector<float> init_params(uint64_t seed) {
  std::vector<float> p(kMlpParams, 0.0f);
  uint64_t s = seed ^ 0x5117A11CE5EEDULL;
  for (int i = 0; i < kMlpHidden * kMlpIn; ++i)
    p[kMlpW1 + i] = static_cast<float>(next_normal(s) * std::sqrt(2.0 / kMlpFeatures));
  for (int i = 0; i < kMlpHidden * kMlpHidden; ++i)
    p[kMlpW2 + i] = static_cast<float>(next_normal(s) * std::sqrt(2.0 / kMlpHidden));
  for (int i = 0; i < kMlpHidden; ++i)
    p[kMlpW3 + i] = static_cast<float>(next_normal(s) * 0.01);
  return p;
}

}  // namespace

MlpDev Context::WarmNet::dev() const {
  MlpDev d;
  d.w1 = w1.get();
  d.w2 = w2.get();
  d.w3 = w3.get();
  d.params = params.get();
  return d;
}

void Context::warm_alloc() {
  ensure_device();
  WarmNet& W = warm_;
  if (W.params.size() == static_cast<size_t>(kMlpParams)) return;
  W.params.resize(kMlpParams);
  W.m.resize(kMlpParams);
  W.v.resize(kMlpParams);
  W.grad.resize(kMlpParams);
  W.w1.resize(kMlpHidden * kMlpIn);
  W.w2.resize(kMlpHidden * kMlpHidden);
  W.w3.resize(kMlpHidden);
  W.loss.resize(1);
  W.m.zero(stream_);
  W.v.zero(stream_);
}

void Context::warm_set(const float* params) {
  warm_alloc();
  WarmNet& W = warm_;
  hip_check(hipMemcpyAsync(W.params.get(), params, kMlpParams * sizeof(float),
                           hipMemcpyHostToDevice, stream_),
            "warm net upload");
  W.m.zero(stream_);
  W.v.zero(stream_);
  W.adam_t = 0;
  launch_mlp_convert(stream_, W.params.get(), W.w1.get(), W.w2.get(), W.w3.get());
  hip_check(hipStreamSynchronize(stream_), "warm net upload");
  W.loaded = true;
}

void Context::warm_get(float* params) {
  hip_check(hipMemcpyAsync(params, warm_.params.get(), kMlpParams * sizeof(float),
                           hipMemcpyDeviceToHost, stream_),
            "warm net download");
  hip_check(hipStreamSynchronize(stream_), "warm net download");
}

void Context::warm_record(int64_t capacity, float rate) {
  ensure_device();
  WarmNet& W = warm_;
  if (capacity <= 0) {
    W.ds_feat.release();
    W.ds_target.release();
    W.ds_count.release();
    W.ds_cap = 0;
    return;
  }
  W.ds_feat.resize(static_cast<size_t>(capacity) * kMlpIn);
  W.ds_target.resize(static_cast<size_t>(capacity));
  W.ds_count.resize(1);
  W.ds_count.zero(stream_);
  hip_check(hipStreamSynchronize(stream_), "record reset");
  W.ds_cap = capacity;
  W.ds_rate = rate;
}

int64_t Context::warm_samples() {
  if (warm_.ds_cap <= 0) return 0;
  unsigned long long n = 0;
  hip_check(hipMemcpyAsync(&n, warm_.ds_count.get(), sizeof(n), hipMemcpyDeviceToHost,
                           stream_),
            "sample count");
  hip_check(hipStreamSynchronize(stream_), "sample count");
  return std::min<int64_t>(static_cast<int64_t>(n), warm_.ds_cap);
}

void Context::warm_train(int steps, int batch, float lr, uint64_t seed, float* loss) {
  const int64_t n = warm_samples();
  if (n <= 0) throw HipError("no recorded samples to train on");
  if (!warm_.loaded) {
    const std::vector<float> p = init_params(seed);
    warm_set(p.data());
  }
  WarmNet& W = warm_;
  const int waves = mlp_train_waves(batch);
  W.grad_part.reserve(static_cast<size_t>(waves) * kMlpParams);
  W.loss_part.reserve(static_cast<size_t>(waves));
  DevBuf<double> losses(std::max(1, steps));
  MlpSamples data;
  data.feat = W.ds_feat.get();
  data.target = W.ds_target.get();
  data.n = n;
  const MlpDev net = W.dev();
  const int64_t salt = static_cast<int64_t>(seed & 0xffffffffu) * 1000003;
  for (int k = 0; k < steps; ++k) {
    launch_mlp_train_pass(stream_, net, data, batch, salt + k, W.grad_part.get(),
                          W.loss_part.get(), waves, nullptr);
    launch_mlp_reduce_adam(stream_, W.grad_part.get(), W.loss_part.get(), waves,
                           W.grad.get(), losses.get() + k, W.params.get(), W.m.get(),
                           W.v.get(), lr, ++W.adam_t);
    launch_mlp_convert(stream_, W.params.get(), W.w1.get(), W.w2.get(), W.w3.get());
  }
  std::vector<double> h(std::max(1, steps));
  hip_check(hipMemcpyAsync(h.data(), losses.get(), h.size() * sizeof(double),
                           hipMemcpyDeviceToHost, stream_),
            "train losses");
  hip_check(hipStreamSynchronize(stream_), "train");
  if (loss != nullptr)
    for (int k = 0; k < steps; ++k) loss[k] = static_cast<float>(h[k]);
}

void Context::warm_debug(int64_t n, const float* feat, const float* target, float* y,
                         float* grad, double* loss) {
  WarmNet& W = warm_;
  // Features rounded to BF16 (nearest even) on the host; the kernels read
  // BF16 directly.
  std::vector<unsigned short> hf(static_cast<size_t>(n) * kMlpIn);
  for (size_t i = 0; i < hf.size(); ++i) {
    uint32_t u;
    std::memcpy(&u, &feat[i], sizeof(u));
    u += 0x7fffu + ((u >> 16) & 1u);
    hf[i] = static_cast<unsigned short>(u >> 16);
  }
  std::vector<float> ht(target, target + n);
  DevBuf<unsigned short> df;
  DevBuf<float> dt, dy;
  df.upload(hf, stream_);
  dt.upload(ht, stream_);
  dy.resize(std::max<int64_t>(1, n));
  const int waves = mlp_train_waves(n);
  W.grad_part.reserve(static_cast<size_t>(waves) * kMlpParams);
  W.loss_part.reserve(static_cast<size_t>(waves));
  MlpSamples data;
  data.feat = df.get();
  data.target = dt.get();
  data.n = n;
  launch_mlp_train_pass(stream_, W.dev(), data, n, -1, W.grad_part.get(),
                        W.loss_part.get(), waves, dy.get());
  launch_mlp_reduce_adam(stream_, W.grad_part.get(), W.loss_part.get(), waves,
                         W.grad.get(), W.loss.get(), W.params.get(), W.m.get(),
                         W.v.get(), 0.0f, 1);
  hip_check(hipMemcpyAsync(y, dy.get(), n * sizeof(float), hipMemcpyDeviceToHost, stream_),
            "debug y");
  hip_check(hipMemcpyAsync(grad, W.grad.get(), kMlpParams * sizeof(float),
                           hipMemcpyDeviceToHost, stream_),
            "debug grad");
  hip_check(hipMemcpyAsync(loss, W.loss.get(), sizeof(double), hipMemcpyDeviceToHost,
                           stream_),
            "debug loss");
  hip_check(hipStreamSynchronize(stream_), "debug net");
}

}  // namespace silk::amd
