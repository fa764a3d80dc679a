// Shared host/device definitions for the gfx950 HF/GW library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace ggpu {

// Library error carrying a status code; the C ABI turns it into a return code
// plus a thread-local message (the reference throws std::runtime_error /
// std::logic_error from the same places).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
  kOk = 0,
  kErrArgument = 1,
  kErrDevice = 2,
  kErrMemory = 3,
  kErrNotPositiveDefinite = 4,
  kErrState = 5,
  kErrSingular = 6,
};

#define GGPU_HIP(call)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw ::ggpu::Error(e_ == hipErrorOutOfMemory ? ::ggpu::kErrMemory : ::ggpu::kErrDevice, \
                          std::string(#call) + ": " + hipGetErrorString(e_));              \
  } while (0)

inline void require(bool ok, const std::string& msg, int code = kErrArgument) {
  if (!ok) throw Error(code, msg);
}

// Complex numbers are (re, im) pairs in memory, like std::complex<T>.
using z64 = double2;
using c32 = float2;

// Owning device allocation.
template <class T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { resize(n); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) GGPU_HIP(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)));
    n_ = n;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

}  // namespace ggpu
