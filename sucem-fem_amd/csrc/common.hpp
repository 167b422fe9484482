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

// Device workspace of a context: numbered slots that keep their allocation from call to
// call and grow when a call needs more, so repeated solves on one context allocate (and
// free, which synchronises the device) nothing after the first.  A slot is used by one
// routine at a time: the Krylov solver takes slots [0, 32), the eigen solver [32, 48) (it
// calls the Krylov solver, so their ranges differ), the batched solver [48, 80).  Calls on
// one context run one at a time (one host thread per context).
struct Workspace {
  std::vector<DBuf<unsigned char>> slot;
  // device memory of at least `bytes` bytes in slot k (nullptr for 0 bytes); the contents
  // are whatever the slot's last user left
  void* get(int k, size_t bytes) {
    if (bytes == 0) return nullptr;
    if ((size_t)k >= slot.size()) slot.resize(k + 1);
    if (slot[k].n < bytes) {
      slot[k].release();
      slot[k].alloc(bytes + (bytes >> 3));   // 1/8 headroom for a slightly larger next call
    }
    return slot[k].p;
  }
  size_t bytes() const {
    size_t t = 0;
    for (const auto& b : slot) t += b.n;
    return t;
  }
  void release() { slot.clear(); }
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int num_cus = 256;
  // pinned host staging for scalar readbacks
  double* pinned = nullptr;
  size_t pinned_n = 0;
  Workspace ws;
  // library-owned stream (non-blocking, default priority) for work captured in hipGraphs:
  // the caller's stream may be the null stream, which cannot be captured
  hipStream_t aux_stream = nullptr;
  hipStream_t graph_stream() {
    if (!aux_stream) SUCEM_HIP(hipStreamCreateWithFlags(&aux_stream, hipStreamNonBlocking));
    return aux_stream;
  }
  void sync() { SUCEM_HIP(hipStreamSynchronize(stream)); }
  void use() { SUCEM_HIP(hipSetDevice(device)); }
  double* host_scratch(size_t n);
};

// typed view of workspace slot k holding `count` elements
template <typename T>
inline T* ws_take(Context* c, int k, int64_t count) {
  return (T*)c->ws.get(k, count > 0 ? sizeof(T) * (size_t)count : 0);
}

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
