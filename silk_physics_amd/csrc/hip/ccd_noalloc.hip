// Earliest TOI of explicit CCD queries without any device allocation: the
// variant of silk_hip_debug_ccd_min for callers that cannot allocate device
// memory (code under HIP graph capture, tight query loops, harnesses that
// forbid hipMalloc).  The queries are staged, and the running minimum kept,
// in a file-scope device array of fixed size; the FP64 CCD kernel is the one
// the step and silk_hip_debug_ccd_min run.
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>

#include "api/context.hpp"
#include "hip/collision_kernels.hpp"
#include "silk_hip.h"

namespace silk::amd {
namespace {

// Layout of the scratch array: the minimum word, then per query its kind
// (int), both segment ends (12 doubles each) and its error filter (3).
constexpr size_t kMinOff = 0;
constexpr size_t kVfOff = 256;
constexpr size_t kP0Off = kVfOff + 4 * static_cast<size_t>(SILK_HIP_CCD_NOALLOC_MAX);
constexpr size_t kP1Off = kP0Off + 96 * static_cast<size_t>(SILK_HIP_CCD_NOALLOC_MAX);
constexpr size_t kErrOff = kP1Off + 96 * static_cast<size_t>(SILK_HIP_CCD_NOALLOC_MAX);
constexpr size_t kScratchBytes = kErrOff + 24 * static_cast<size_t>(SILK_HIP_CCD_NOALLOC_MAX);
static_assert(kP0Off % 8 == 0 && kP1Off % 8 == 0 && kErrOff % 8 == 0, "alignment");

// Each device that runs the entry point holds its own copy until the
// process ends.  In host code this name is not a device address: only the
// pointer hipGetSymbolAddress returns for the current device is.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
std::array<unsigned char*, kMaxDevices> g_ptr{};

// The device address of `scratch` on `device`, looked up once per device
// with that device current.
unsigned char* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) throw HipError("device id out of range");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_ptr[device] == nullptr) {
    hip_check(hipSetDevice(device), "hipSetDevice");
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (p == nullptr) throw HipError("hipGetSymbolAddress(scratch) returned null");
    g_ptr[device] = static_cast<unsigned char*>(p);
  }
  return g_ptr[device];
}

}  // namespace

int ccd_min_noalloc(Context& c, int64_t n, const int32_t* is_vf, const double* p0,
                    const double* p1, const double* err, double ms, double tolerance,
                    int32_t max_iter, double* min_toi) {
  if (n < 0 || min_toi == nullptr) {
    c.last_error_ = "ccd_min_noalloc: n < 0 or min_toi is null";
    return SILK_HIP_ERR_INVALID_CONFIG;
  }
  if (n > SILK_HIP_CCD_NOALLOC_MAX) {
    c.last_error_ = "ccd_min_noalloc: " + std::to_string(n) +
                    " queries exceed the fixed scratch capacity of " +
                    std::to_string(SILK_HIP_CCD_NOALLOC_MAX);
    return SILK_HIP_ERR_INVALID_CONFIG;
  }
  *min_toi = 1.0;
  if (n == 0) return SILK_HIP_OK;
  if (!is_vf || !p0 || !p1 || !err) {
    c.last_error_ = "ccd_min_noalloc: null query array";
    return SILK_HIP_ERR_INVALID_CONFIG;
  }
  c.ensure_device();
  unsigned char* base = scratch_on(c.device_);
  const size_t m = static_cast<size_t>(n);
  hipStream_t s = c.stream_;
  // The minimum starts at 1 (no hit) on every call: the array keeps the last
  // call's contents.
  const double one = 1.0;
  hip_check(hipMemcpyAsync(base + kMinOff, &one, sizeof(double), hipMemcpyHostToDevice, s),
            "noalloc min init");
  hip_check(hipMemcpyAsync(base + kVfOff, is_vf, 4 * m, hipMemcpyHostToDevice, s),
            "noalloc kinds");
  hip_check(hipMemcpyAsync(base + kP0Off, p0, 96 * m, hipMemcpyHostToDevice, s),
            "noalloc t0");
  hip_check(hipMemcpyAsync(base + kP1Off, p1, 96 * m, hipMemcpyHostToDevice, s),
            "noalloc t1");
  hip_check(hipMemcpyAsync(base + kErrOff, err, 24 * m, hipMemcpyHostToDevice, s),
            "noalloc err");
  launch_ccd_min(s, static_cast<int>(n), reinterpret_cast<const int*>(base + kVfOff),
                 reinterpret_cast<const double*>(base + kP0Off),
                 reinterpret_cast<const double*>(base + kP1Off),
                 reinterpret_cast<const double*>(base + kErrOff), ms, tolerance, max_iter,
                 reinterpret_cast<unsigned long long*>(base + kMinOff), c.ccd_zero_mode());
  hip_check(hipGetLastError(), "noalloc ccd launch");
  hip_check(hipMemcpyAsync(min_toi, base + kMinOff, sizeof(double), hipMemcpyDeviceToHost, s),
            "noalloc min");
  hip_check(hipStreamSynchronize(s), "noalloc sync");
  return SILK_HIP_OK;
}

}  // namespace silk::amd
