// Element products of the matrix-free operators with the reference tensors folded into
// the instructions (shared by the per-cell and the patch kernels).
#pragma once
#include <utility>

#include "internal.hpp"
#include "kernels.hpp"
#include "mf_tables.inc"

namespace sucem {

// ---- element matrix with the reference tensors folded into the instructions ----------
// kMfNum<order>[k][p] / kMfDen<order>: k < 6 mass M_k, k >= 6 curl-curl C_{k-6}; p runs
// over the upper triangle (i <= j) row by row.  The numerators are small integers
// (exact instruction literals in FP32 and FP64); zero entries (~40 % for order 2)
// generate no code.
template <int ORDER>
struct MfTab;
template <>
struct MfTab<1> {
  static constexpr int NLOC = 6, NP = kMfNp1;
  static constexpr double den = kMfDen1;
  static constexpr double v(int k, int p) { return kMfNum1[k][p]; }
};
template <>
struct MfTab<2> {
  static constexpr int NLOC = 20, NP = kMfNp2;
  static constexpr double den = kMfDen2;
  static constexpr double v(int k, int p) { return kMfNum2[k][p]; }
};

// (i, j) of upper-triangle pair p
template <int NLOC>
constexpr int pair_i(int p) {
  int i = 0;
  while (p >= NLOC - i) { p -= NLOC - i; ++i; }
  return i;
}
template <int NLOC>
constexpr int pair_j(int p) {
  int i = 0;
  while (p >= NLOC - i) { p -= NLOC - i; ++i; }
  return i + p;
}

template <typename R, int ORDER, int P, int K>
__device__ __forceinline__ void k_term(R& kij, const R* a) {
  constexpr double v = MfTab<ORDER>::v(K, P);
  if constexpr (v != 0.0) kij += a[K] * (R)v;
}

template <int ORDER, int P, int... K>
constexpr bool pair_nonzero(std::integer_sequence<int, K...>) {
  return ((MfTab<ORDER>::v(K, P) != 0.0) || ...);
}

template <typename R, int ORDER, int P, int... K>
__device__ __forceinline__ void pair_step(const R* a, const typename C2<R>::T* xl,
                                          typename C2<R>::T* acc,
                                          std::integer_sequence<int, K...> ks) {
  if constexpr (pair_nonzero<ORDER, P>(ks)) {
    constexpr int N = MfTab<ORDER>::NLOC;
    constexpr int i = pair_i<N>(P), j = pair_j<N>(P);
    R kij = R(0);
    (k_term<R, ORDER, P, K>(kij, a), ...);
    acc[i].x += kij * xl[j].x;
    acc[i].y += kij * xl[j].y;
    if constexpr (i != j) {
      acc[j].x += kij * xl[i].x;
      acc[j].y += kij * xl[i].y;
    }
  }
}

template <typename R, int ORDER, int... P>
__device__ __forceinline__ void element_product(const R* a, const typename C2<R>::T* xl,
                                                typename C2<R>::T* acc,
                                                std::integer_sequence<int, P...>) {
  (pair_step<R, ORDER, P>(a, xl, acc, std::make_integer_sequence<int, 12>{}), ...);
}

// Half of the rows at a time (HALF 0: rows i < N/2, HALF 1: rows >= N/2): only that
// half's accumulators are live, which lowers the register count (more waves per SIMD)
// at the price of forming the K entries that couple the halves twice.
template <typename R, int ORDER, int HALF, int P, int... K>
__device__ __forceinline__ void pair_step_half(const R* a, const typename C2<R>::T* xl,
                                               typename C2<R>::T* acc,
                                               std::integer_sequence<int, K...> ks) {
  constexpr int N = MfTab<ORDER>::NLOC, H = N / 2;
  constexpr int i = pair_i<N>(P), j = pair_j<N>(P);
  constexpr bool mine = HALF == 0 ? (i < H) : (j >= H);
  if constexpr (mine && pair_nonzero<ORDER, P>(ks)) {
    R kij = R(0);
    (k_term<R, ORDER, P, K>(kij, a), ...);
    if constexpr (HALF == 0) {
      acc[i].x += kij * xl[j].x;
      acc[i].y += kij * xl[j].y;
      if constexpr (j < H && j != i) {
        acc[j].x += kij * xl[i].x;
        acc[j].y += kij * xl[i].y;
      }
    } else {
      acc[j - H].x += kij * xl[i].x;
      acc[j - H].y += kij * xl[i].y;
      if constexpr (i >= H && i != j) {
        acc[i - H].x += kij * xl[j].x;
        acc[i - H].y += kij * xl[j].y;
      }
    }
  }
}

template <typename R, int ORDER, int HALF, int... P>
__device__ __forceinline__ void element_product_half(const R* a, const typename C2<R>::T* xl,
                                                     typename C2<R>::T* acc,
                                                     std::integer_sequence<int, P...>) {
  (pair_step_half<R, ORDER, HALF, P>(a, xl, acc, std::make_integer_sequence<int, 12>{}), ...);
}

}  // namespace sucem
