// Device helpers shared by the operator and Krylov translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace sucem {

template <typename R> struct C2;
template <> struct C2<float> { using T = float2; };
template <> struct C2<double> { using T = double2; };

__device__ __forceinline__ float2 mk(float a, float b) { return make_float2(a, b); }
__device__ __forceinline__ double2 mk(double a, double b) { return make_double2(a, b); }

template <typename T>
__device__ __forceinline__ T cmul(T a, T b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 to_d(float2 a) { return make_double2(a.x, a.y); }
__device__ __forceinline__ double2 to_d(double2 a) { return a; }

// Wavefront-then-block sum of a double2; every thread of the block gets the result.
// Deterministic: fixed reduction tree.
template <int TB>
__device__ __forceinline__ double2 block_sum(double2 v) {
  __shared__ double2 red[TB / 64];
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o);
    v.y += __shfl_xor(v.y, o);
  }
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double2 s = red[0];
  for (int i = 1; i < TB / 64; ++i) { s.x += red[i].x; s.y += red[i].y; }
  return s;
}

// Sum of G per-block partials, computed redundantly (and identically) by every block.
template <int TB>
__device__ __forceinline__ double2 sum_partials(const double2* p, int G) {
  double2 v = make_double2(0.0, 0.0);
  for (int i = threadIdx.x; i < G; i += TB) { v.x += p[i].x; v.y += p[i].y; }
  return block_sum<TB>(v);
}

// Logical block index that gives each XCD a contiguous share of the rows
// (blocks are dispatched round-robin over the 8 XCDs).
__device__ __forceinline__ int xcd_block(int G) {
  const int b = blockIdx.x;
  return (b & 7) * (G >> 3) + (b >> 3);
}

}  // namespace sucem
