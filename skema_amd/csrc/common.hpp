// Shared host/device definitions for libskema_mi355x.
// Column-major (LayoutLeft) FP64 storage throughout, like the reference's
// Kokkos views (src/Skema_Utils.hpp:9-18).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace skema {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Status codes mirrored in include/skema_mi355x.h.
enum Status : int {
  kOk = 0,
  kErrArg = 1,
  kErrShape = 2,
  kErrHip = 3,
  kErrOom = 4,
  kErrNumeric = 5,
  kErrState = 6,
  kErrUnsupported = 7,
};

// Library error carried up to the C ABI, which turns it into a status code
// plus a thread-local message.  Nothing throws across the ABI.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SKEMA_HIP(expr)                                                        \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      throw ::skema::Error(e_ == hipErrorOutOfMemory ? ::skema::kErrOom        \
                                                     : ::skema::kErrHip,       \
                           std::string(#expr) + ": " + hipGetErrorString(e_)); \
    }                                                                          \
  } while (0)

#define SKEMA_REQUIRE(cond, code, msg)                  \
  do {                                                  \
    if (!(cond)) throw ::skema::Error((code), (msg));   \
  } while (0)

inline void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw Error(kErrHip, std::string("launch ") + what + ": " + hipGetErrorString(e));
}

// Device buffer owned by a library object.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t b) {
    release();
    if (b == 0) return;
    SKEMA_HIP(hipMalloc(&p, b));
    bytes = b;
  }
  // Grow-only scratch.
  void reserve(size_t b) { if (b > bytes) alloc(b); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

__host__ __device__ inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace skema
