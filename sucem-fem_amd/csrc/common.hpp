// Shared internals of libsucem_hip: error state, device buffers, context.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "sucem_hip.h"

namespace sucem {

// Error carrying a C-ABI status code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define SUCEM_HIP(call)                                                           \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      throw ::sucem::Error(e_ == hipErrorOutOfMemory ? SUCEM_ERR_OOM : SUCEM_ERR_HIP, \
                           std::string(#call) + ": " + hipGetErrorString(e_) +     \
                               " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
    }                                                                             \
  } while (0)

#define SUCEM_CHECK(cond, code, msg)                    \
  do {                                                  \
    if (!(cond)) throw ::sucem::Error((code), (msg));   \
  } while (0)

// Owning device allocation.
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) SUCEM_HIP(hipMalloc(&p, count * sizeof(T)));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int num_cus = 256;
  // pinned host staging for scalar readbacks
  double* pinned = nullptr;
  size_t pinned_n = 0;
  void sync() { SUCEM_HIP(hipStreamSynchronize(stream)); }
  void use() { SUCEM_HIP(hipSetDevice(device)); }
  double* host_scratch(size_t n);
};

template <typename T>
inline void h2d(Context* c, T* d, const T* h, size_t n) {
  if (n) SUCEM_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
}
template <typename T>
inline void d2h(Context* c, T* h, const T* d, size_t n) {
  if (n) SUCEM_HIP(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
}

inline unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

#define SUCEM_LAUNCH(kernel, grid, block, shmem, stream, ...)                          \
  do {                                                                                 \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shmem, stream, __VA_ARGS__);  \
    SUCEM_HIP(hipGetLastError());                                                      \
  } while (0)

}  // namespace sucem

// Opaque handle definitions (C-ABI types).
struct sucem_ctx {
  sucem::Context c;
};
