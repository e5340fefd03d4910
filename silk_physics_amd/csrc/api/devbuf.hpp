// Owning device buffer and HIP error helpers for the runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

namespace silk::amd {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw HipError(std::string(what) + ": " + hipGetErrorString(e));
  }
}

template <typename T>
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t n) { resize(n); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }

  void resize(size_t n) {
    release();
    if (n > 0) hip_check(hipMalloc(&p_, n * sizeof(T)), "hipMalloc");
    n_ = n;
  }
  /// Grow-only reallocation (contents are not preserved).
  void reserve(size_t n) {
    if (n > n_) resize(n);
  }
  void release() {
    if (p_ != nullptr) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  void upload(const std::vector<T>& h, hipStream_t s) {
    resize(h.size());
    if (!h.empty())
      hip_check(hipMemcpyAsync(p_, h.data(), h.size() * sizeof(T),
                               hipMemcpyHostToDevice, s),
                "upload");
  }
  void zero(hipStream_t s) {
    if (n_ > 0)
      hip_check(hipMemsetAsync(p_, 0, n_ * sizeof(T), s), "memset");
  }
  T* get() { return p_; }
  const T* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

}  // namespace silk::amd
