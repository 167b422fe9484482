// Matrix-free fused operator y = (cs S + ct T + i c0 S0) x.
//
// The assembled operator streams one (S, T) pair and one column index per nonzero:
// 12 B (FP32) or 20 B (FP64) per nonzero, ~43 nonzeros per order-2 row.  Here the
// volume part is recomputed from 12 geometric coefficients per cell instead:
//
//  1. k_mf_cells: one thread per cell forms the dense element matrix
//     K = sum_k (ct cT_k) M_k + (cs cS_k) C_k (symmetric, nloc x nloc) with the
//     reference tensors M_k, C_k compiled into the instructions (generated tables;
//     zero entries generate nothing), and writes
//     the element product K x_c (nloc complex values) to a per-cell buffer;
//  2. k_mf_faces: one thread per boundary face forms its ABC matrix i c0 S0_f
//     (nfd x nfd) and writes S0_f x_f the same way;
//  3. k_mf_gather: each DOF row sums its products -- they are stored in DOF-sorted
//     order (cells first, then faces, each in index order: a permutation computed
//     once per space), so a row is one contiguous segment -- then applies the PEC row
//     mask and the optional Krylov epilogue.
//
// HBM traffic per application: 12 coefficients, nloc DOF ids and nloc product positions
// per cell, x, the product buffer (written and read once), the segment pointers, y --
// about a fifth of the assembled operator's bytes for order 2 -- for ~2.1k FMAs per
// order-2 cell.  The order of every sum is
// fixed, so results are deterministic.  The element buffer belongs to the System:
// concurrent applications of one System on different streams are not supported.
#include <algorithm>
#include <cstring>
#include <vector>
#include <utility>
#include <rocprim/rocprim.hpp>

#include "geom.hpp"
#include "internal.hpp"
#include "kernels.hpp"
#include "ops.hpp"
#include "mf_element.hpp"

namespace sucem {
namespace {

constexpr int TB = 256;

__constant__ int c_msym[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};

// Per cell, SoA: coef[k * nc + c], k < 6: eps_r |det J| (J^-1 J^-T)_k (pairs with M_k),
// k >= 6: (1/mu_r) / |det J| (J^T J)_{k-6} (pairs with C_{k-6}).
template <typename R>
__global__ void k_mf_coefs(int64_t nc, const double* xyz, const int32_t* cells, const double* eps,
                           const double* imu, R* coef) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc;
       c += (int64_t)gridDim.x * blockDim.x) {
    double J[3][3], Ji[3][3], det;
    cell_jacobian(xyz, cells + 4 * c, J, det, Ji);
    const double adet = fabs(det);
    const double e = eps[c] * adet, m = imu[c] / adet;
    for (int k = 0; k < 6; ++k) {
      const int a = c_msym[k][0], b = c_msym[k][1];
      const double g = Ji[a][0] * Ji[b][0] + Ji[a][1] * Ji[b][1] + Ji[a][2] * Ji[b][2];
      const double q = J[0][a] * J[0][b] + J[1][a] * J[1][b] + J[2][a] * J[2][b];
      coef[k * nc + c] = (R)(e * g);
      coef[(6 + k) * nc + c] = (R)(m * q);
    }
  }
}

__global__ void k_mf_transpose_dofs(int64_t nc, int nloc, const int32_t* cd, int32_t* cdT) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nc * nloc;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = i / nloc;
    const int l = (int)(i % nloc);
    cdT[(int64_t)l * nc + c] = cd[i];
  }
}

__global__ void k_mf_iota(int64_t n, uint32_t* v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] = (uint32_t)i;
}

// Product slots: cells s < nc * nloc (s = c * nloc + l), then faces (nc * nloc + f * nfd + a).
// posT[l * nc + c] / posF[f * nfd + a] = the slot's position in the DOF-sorted order.
__global__ void k_mf_inverse(int64_t n, int64_t nc, int nloc, const uint32_t* slot, int32_t* posT,
                             int32_t* posF) {
  const int64_t ncl = nc * nloc;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = slot[j];
    if (s < ncl) posT[(s % nloc) * nc + s / nloc] = (int32_t)j;
    else posF[s - ncl] = (int32_t)j;
  }
}

// Sort keys: the DOF of every cell slot, then of every boundary-face slot.
__global__ void k_mf_keys(int64_t nc, int nloc, int64_t nbf, int nfd, const int32_t* cd,
                          const int32_t* bf_cell, const int32_t* bf_local,
                          const int32_t* face_dofs, uint32_t* keys) {
  const int64_t ncl = nc * nloc, n = ncl + nbf * nfd;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i < ncl) {
      keys[i] = (uint32_t)cd[i];
    } else {
      const int64_t f = (i - ncl) / nfd;
      const int a = (int)((i - ncl) % nfd);
      keys[i] = (uint32_t)cd[(int64_t)bf_cell[f] * nloc + face_dofs[bf_local[f] * nfd + a]];
    }
  }
}

__global__ void k_mf_soa6(int64_t nbf, const double* H, double* h64, float* h32) {
  for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nbf;
       f += (int64_t)gridDim.x * blockDim.x)
    for (int k = 0; k < 6; ++k) {
      h64[k * nbf + f] = H[6 * f + k];
      h32[k * nbf + f] = (float)H[6 * f + k];
    }
}

// ABC face products: v = i c0 S0_f x_f written to the face slots (zero when c0 == 0).
// One thread per (face, face-DOF row): 8 threads share a face's coefficients and x
// values through the cache; the reference face tensors are read from LDS.
template <typename R, int NFD>
__global__ __launch_bounds__(TB) void k_mf_faces(int64_t nbf, int nloc, const R* __restrict__ h,
                                                 const int32_t* __restrict__ bf_cell,
                                                 const int32_t* __restrict__ bf_local,
                                                 const int32_t* __restrict__ cd,
                                                 const int32_t* __restrict__ face_dofs,
                                                 const double* __restrict__ ref_face, R c0,
                                                 const int32_t* __restrict__ posF,
                                                 const typename C2<R>::T* __restrict__ x,
                                                 typename C2<R>::T* __restrict__ yl) {
  using C = typename C2<R>::T;
  constexpr int KK = NFD * NFD;
  __shared__ R rf_s[4 * 6 * KK];
  for (int t = threadIdx.x; t < 4 * 6 * KK; t += TB) rf_s[t] = (R)ref_face[t];
  __syncthreads();
  const int64_t total = nbf * NFD;
  for (int64_t id = blockIdx.x * (int64_t)TB + threadIdx.x; id < total;
       id += (int64_t)gridDim.x * TB) {
    const int64_t f = id / NFD;
    const int a = (int)(id % NFD);
    const int lf = bf_local[f];
    const int64_t c = bf_cell[f];
    R hk[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) hk[k] = h[k * nbf + f];
    const R* rf = rf_s + lf * 6 * KK + a * NFD;
    C v = mk(R(0), R(0));
#pragma unroll
    for (int b = 0; b < NFD; ++b) {
      R m = R(0);
#pragma unroll
      for (int k = 0; k < 6; ++k) m += hk[k] * rf[k * KK + b];
      const C xb = x[cd[c * nloc + face_dofs[lf * NFD + b]]];
      v.x += m * xb.x;
      v.y += m * xb.y;
    }
    yl[posF[id]] = mk(-c0 * v.y, c0 * v.x);
  }
}

// dptr[d] = first position of DOF d in the sorted key list (every DOF occurs).
__global__ void k_mf_heads(int64_t n, int64_t ndofs, const uint32_t* keys, int32_t* dptr) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i == 0 || keys[i] != keys[i - 1]) dptr[keys[i]] = (int32_t)i;
    if (i == n - 1) dptr[ndofs] = (int32_t)n;
  }
}

// Element products, one thread per cell; consecutive blocks of cells go to one XCD.
// With the half-row split the registers are capped so that more waves share a SIMD
// (3 FP64, 5 FP32) and the loads of one wave overlap the FMAs of the others; the
// autotuner compares it with the uncapped whole-element form.
template <typename R, int ORDER, bool SPLIT>
__global__ __launch_bounds__(TB, SPLIT ? (sizeof(R) == 8 ? 3 : 5) : 1) void k_mf_cells(
    int64_t nc, int G, const R* __restrict__ coef,
                                                 const int32_t* __restrict__ cdT,
                                                 const int32_t* __restrict__ posT, R cs, R ct,
                                                 const typename C2<R>::T* __restrict__ x,
                                                 typename C2<R>::T* __restrict__ yl) {
  using C = typename C2<R>::T;
  constexpr int NLOC = MfTab<ORDER>::NLOC;
  // one cell per thread and no loop, so that no constant is hoisted into a register
  const int64_t c = (int64_t)xcd_block(G) * TB + threadIdx.x;
  if (c < nc) {
    R a[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) a[k] = coef[k * nc + c] * (k < 6 ? ct : cs) * R(1.0 / MfTab<ORDER>::den);
    C xl[NLOC];
#pragma unroll
    for (int i = 0; i < NLOC; ++i) xl[i] = x[cdT[(int64_t)i * nc + c]];
    // each product goes to its place in the DOF-sorted order: the contributions of
    // neighbouring cells to one DOF land side by side, and the gather streams them
    if constexpr (!SPLIT) {
      C acc[NLOC];
#pragma unroll
      for (int i = 0; i < NLOC; ++i) acc[i] = mk(R(0), R(0));
      element_product<R, ORDER>(a, xl, acc, std::make_integer_sequence<int, MfTab<ORDER>::NP>{});
#pragma unroll
      for (int i = 0; i < NLOC; ++i) yl[posT[(int64_t)i * nc + c]] = acc[i];
    } else {
      constexpr int H = NLOC / 2;
      C acc[H];
#pragma unroll
      for (int i = 0; i < H; ++i) acc[i] = mk(R(0), R(0));
      element_product_half<R, ORDER, 0>(a, xl, acc,
                                        std::make_integer_sequence<int, MfTab<ORDER>::NP>{});
#pragma unroll
      for (int i = 0; i < H; ++i) {
        yl[posT[(int64_t)i * nc + c]] = acc[i];
        acc[i] = mk(R(0), R(0));
      }
      // opaque to the optimiser: the coupling entries are formed again rather than kept
      // live across the halves (which would cost more registers than the split saves)
#pragma unroll
      for (int k = 0; k < 12; ++k) asm volatile("" : "+v"(a[k]));
      element_product_half<R, ORDER, 1>(a, xl, acc,
                                        std::make_integer_sequence<int, MfTab<ORDER>::NP>{});
#pragma unroll
      for (int i = 0; i < H; ++i) yl[posT[(int64_t)(H + i) * nc + c]] = acc[i];
    }
  }
}

// ---- order-2 element products on the matrix cores (FP32 operator) ---------------------
// Z = M X with M the 12 reference tensors' integer numerators stacked as 240 rows (tensor
// k, row i) x 20 columns j (K padded to 32), exact in fp16 (|v| <= 2268, at most 11
// significant bits), and X 8 cells' x, real and imaginary parts as 16 columns.  Each column
// is scaled by a power of two s (its largest |x| into [1, 2)) and split into fp16 pieces
// hi + 2^-11 lo (22 significant bits); v_mfma_f32_16x16x32_f16 forms M lo, which is
// scaled by 2^-11, then accumulates M hi onto it, per 16-row tile.  The contraction
// y_i = sum_k a_k Z[(k, i)] / s runs on the vector units: M's rows are ordered so that the
// 4 output rows a lane holds per tile (rows 4 (lane / 16) .. + 3), over the 15 tiles, are
// 5 rows i of all 12 tensors k, so no sums cross lanes.  MFMA operand layout (checked on
// the device): A[row = lane % 16][k = 8 (lane / 16) + e], B[k = 8 (lane / 16) + e][col =
// lane % 16], D[4 (lane / 16) + q][lane % 16].
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));
constexpr int kMfmaTiles = 15;   // 240 rows / 16

// (i, k) held at slot = 4 t + q by lane group g: i = 5 g + slot / 12, k = slot % 12
__host__ __device__ constexpr int mfma_row_i(int g, int slot) { return 5 * g + slot / 12; }
__host__ __device__ constexpr int mfma_row_k(int slot) { return slot % 12; }

// A block takes 32 cells at a time (4 waves, one tile of 8 cells each): their 12
// coefficients, 20 x values, places and products go through LDS with coalesced,
// once-per-cell loads and stores; the waves read their MFMA operands from there.
constexpr int kMfmaCells = 32;   // cells per block and step (TB / 64 tiles of 8)
__global__ __launch_bounds__(TB) void k_mf_cells_mfma(int64_t nc, const float* __restrict__ coef,
                                                      const int32_t* __restrict__ cdT,
                                                      const int32_t* __restrict__ posT, float cs,
                                                      float ct, const half8_t* __restrict__ afrag,
                                                      const float2* __restrict__ x,
                                                      float2* __restrict__ yl) {
  static_assert(TB == 256 && kMfmaCells == 8 * (TB / 64), "one tile of 8 cells per wave");
  __shared__ float s_a[12][kMfmaCells];
  __shared__ float2 s_x[20][kMfmaCells];
  __shared__ float2 s_y[20][kMfmaCells];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, col = lane & 15, part = col & 1;
  const int cl = 8 * wave + (col >> 1);            // this lane's cell in the block's 32
  half8_t A[kMfmaTiles];
#pragma unroll
  for (int t = 0; t < kMfmaTiles; ++t) A[t] = afrag[t * 64 + lane];
  const float inv_den = (float)(1.0 / MfTab<2>::den);
  const int64_t ngroup = (nc + kMfmaCells - 1) / kMfmaCells;
  for (int64_t grp = blockIdx.x; grp < ngroup; grp += gridDim.x) {   // uniform per block
    const int64_t c0 = grp * kMfmaCells;
    for (int e = threadIdx.x; e < 12 * kMfmaCells; e += TB) {
      const int k = e / kMfmaCells, ci = e % kMfmaCells;
      const int64_t c = c0 + ci;
      s_a[k][ci] = c < nc ? coef[k * nc + c] * (k < 6 ? ct : cs) * inv_den : 0.f;
    }
    for (int e = threadIdx.x; e < 20 * kMfmaCells; e += TB) {
      const int j = e / kMfmaCells, ci = e % kMfmaCells;
      const int64_t c = c0 + ci;
      s_x[j][ci] = c < nc ? x[cdT[(int64_t)j * nc + c]] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float xv[8];
    float m = 0.f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int j = 8 * g + e;
      float v = 0.f;
      if (j < 20) {
        const float2 xd = s_x[j][cl];
        v = part ? xd.y : xd.x;
      }
      xv[e] = v;
      m = fmaxf(m, fabsf(v));
    }
    m = fmaxf(m, __shfl_xor(m, 16));
    m = fmaxf(m, __shfl_xor(m, 32));
    int ex = 0;
    (void)frexpf(m, &ex);                          // m = f 2^ex, f in [0.5, 1)
    const float sc = m > 0.f ? ldexpf(1.f, 1 - ex) : 1.f;
    half8_t hi, lo;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float v = xv[e] * sc;
      const _Float16 h = (_Float16)v;
      hi[e] = h;
      lo[e] = (_Float16)((v - (float)h) * 2048.f);
    }
    float a[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) a[k] = s_a[k][cl];
    float y[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < kMfmaTiles; ++t) {
      float4_t acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[t], lo, acc, 0, 0, 0);
      acc = acc * (1.f / 2048.f);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[t], hi, acc, 0, 0, 0);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int slot = 4 * t + q;
        y[slot / 12] += a[mfma_row_k(slot)] * acc[q];
      }
    }
    const float inv = 1.f / sc;
    float* sy = reinterpret_cast<float*>(&s_y[0][0]);
#pragma unroll
    for (int ii = 0; ii < 5; ++ii)
      sy[2 * (mfma_row_i(g, 12 * ii) * kMfmaCells + cl) + part] = y[ii] * inv;
    __syncthreads();
    for (int e = threadIdx.x; e < 20 * kMfmaCells; e += TB) {
      const int i = e / kMfmaCells, ci = e % kMfmaCells;
      const int64_t c = c0 + ci;
      if (c < nc) yl[posT[(int64_t)i * nc + c]] = s_y[i][ci];
    }
    __syncthreads();   // s_a, s_x and s_y are refilled by the next step
  }
}

// fp16 bits of an integer with at most 11 significant bits (exact)
uint16_t fp16_of_int(int v) {
  if (v == 0) return 0;
  const uint16_t sign = v < 0 ? 0x8000 : 0;
  unsigned a = (unsigned)(v < 0 ? -v : v);
  int e = 0;
  while ((a >> (e + 1)) != 0) ++e;                  // a in [2^e, 2^(e+1))
  SUCEM_CHECK(e <= 15 && (e <= 10 || (a & ((1u << (e - 10)) - 1)) == 0), SUCEM_ERR_INVALID,
              "reference tensor entry not exact in fp16");
  const unsigned mant = e >= 10 ? (a >> (e - 10)) & 0x3FF : (a << (10 - e)) & 0x3FF;
  return (uint16_t)(sign | ((unsigned)(e + 15) << 10) | mant);
}

// the MFMA A operand of every tile and lane (see k_mf_cells_mfma), built once per space
void mfma_fragments(Context* c, Space* s) {
  if (s->mf_afrag.p) return;
  constexpr int N = 20;
  auto num = [](int k, int i, int j) {
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    return kMfNum2[k][lo * N - lo * (lo - 1) / 2 + (hi - lo)];
  };
  std::vector<uint16_t> h((size_t)kMfmaTiles * 64 * 8, 0);
  for (int t = 0; t < kMfmaTiles; ++t)
    for (int lane = 0; lane < 64; ++lane) {
      const int r = lane % 16, gr = r / 4, slot = 4 * t + r % 4;
      const int i = mfma_row_i(gr, slot), k = mfma_row_k(slot);
      for (int e = 0; e < 8; ++e) {
        const int j = 8 * (lane / 16) + e;
        h[((size_t)t * 64 + lane) * 8 + e] = j < N ? fp16_of_int(num(k, i, j)) : 0;
      }
    }
  s->mf_afrag.alloc(h.size());
  h2d(c, s->mf_afrag.p, h.data(), h.size());
  c->sync();
}

// Row sums of the DOF-sorted products: L lanes per row, each summing every L-th
// product of the row's contiguous segment (a wave covers one contiguous stretch of the
// buffer), then the PEC row mask and the epilogue.  The summation order is fixed per
// row, so results are deterministic.
template <typename R, int EPI, int L>
__global__ __launch_bounds__(TB) void k_mf_gather(
    int n, int G, const int32_t* __restrict__ dptr, const typename C2<R>::T* __restrict__ yl,
    const uint8_t* __restrict__ mask, typename C2<R>::T* __restrict__ y,
    const typename C2<R>::T* __restrict__ w, RedSite rs) {
  using C = typename C2<R>::T;
  constexpr int RPB = TB / L;
  const int lane = threadIdx.x % L;
  const int lb = xcd_block(G);
  int chunk = (n + G - 1) / G;
  chunk = (chunk + RPB - 1) / RPB * RPB;
  const int start = lb * chunk;
  const int end = min(n, start + chunk);
  double2 d1 = make_double2(0.0, 0.0), d2 = make_double2(0.0, 0.0);
  for (int base = start; base < end; base += RPB) {
    const int r = base + threadIdx.x / L;
    C acc = mk(R(0), R(0));
    if (r < end) {
      const int j1 = dptr[r + 1];
      for (int j = dptr[r] + lane; j < j1; j += L) {
        const C v = yl[j];
        acc.x += v.x;
        acc.y += v.y;
      }
    }
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) {
      acc.x += __shfl_xor(acc.x, o);
      acc.y += __shfl_xor(acc.y, o);
    }
    if (lane == 0 && r < end) {
      if (mask && mask[r]) acc = mk(R(0), R(0));
      y[r] = acc;
      if (EPI) {
        const double2 yd = to_d(acc);
        double2 wd = to_d(w[r]);
        if (EPI == 2) wd.y = -wd.y;
        d1.x += wd.x * yd.x - wd.y * yd.y;
        d1.y += wd.x * yd.y + wd.y * yd.x;
        d2.x += yd.x * yd.x + yd.y * yd.y;
      }
    }
  }
  if (EPI) reduce_finalize<TB>(rs, G, d1, d2);
}

// The same row sums with the block's products staged through LDS: a block's rows own
// one contiguous stretch of the product buffer, which is read fully coalesced in
// pieces of CAP products; every thread then sums its own segment from LDS in index
// order (the same order as the one-lane form, hence the same result).
template <typename R, int EPI>
__global__ __launch_bounds__(TB) void k_mf_gather_lds(
    int n, int G, const int32_t* __restrict__ dptr, const typename C2<R>::T* __restrict__ yl,
    const uint8_t* __restrict__ mask, typename C2<R>::T* __restrict__ y,
    const typename C2<R>::T* __restrict__ w, RedSite rs) {
  using C = typename C2<R>::T;
  constexpr int CAP = 4 * TB;
  __shared__ C buf[CAP];
  const int lb = xcd_block(G);
  int chunk = (n + G - 1) / G;
  chunk = (chunk + TB - 1) / TB * TB;
  const int start = lb * chunk;
  const int end = min(n, start + chunk);
  double2 d1 = make_double2(0.0, 0.0), d2 = make_double2(0.0, 0.0);
  for (int base = start; base < end; base += TB) {
    const int r = base + threadIdx.x;
    const int s0 = r < end ? dptr[r] : 0, s1 = r < end ? dptr[r + 1] : 0;
    const int p0 = dptr[base], p1 = dptr[min(base + TB, end)];
    C acc = mk(R(0), R(0));
    for (int q0 = p0; q0 < p1; q0 += CAP) {
      const int q1 = min(p1, q0 + CAP);
      __syncthreads();
      for (int q = q0 + threadIdx.x; q < q1; q += TB) buf[q - q0] = yl[q];
      __syncthreads();
      const int a = max(s0, q0), b = min(s1, q1);
      for (int q = a; q < b; ++q) {
        const C v = buf[q - q0];
        acc.x += v.x;
        acc.y += v.y;
      }
    }
    if (r < end) {
      if (mask && mask[r]) acc = mk(R(0), R(0));
      y[r] = acc;
      if (EPI) {
        const double2 yd = to_d(acc);
        double2 wd = to_d(w[r]);
        if (EPI == 2) wd.y = -wd.y;
        d1.x += wd.x * yd.x - wd.y * yd.y;
        d1.y += wd.x * yd.y + wd.y * yd.x;
        d2.x += yd.x * yd.x + yd.y * yd.y;
      }
    }
  }
  if (EPI) reduce_finalize<TB>(rs, G, d1, d2);
}

}  // namespace

void mf_prepare(Context* c, System* sys) {
  if (sys->mf_ready) return;
  Space* s = sys->space;
  const Mesh* m = s->mesh;
  const int64_t nc = m->nc, nbf = s->nbf;
  const int nloc = s->nloc, nfd = s->nfd;
  std::lock_guard<std::mutex> hold(s->lazy);
  if (!s->mf_cdT.p) {
    // DOF ids per cell (SoA); products sorted by (DOF, slot) and their positions
    s->mf_cdT.alloc((size_t)nc * nloc);
    SUCEM_LAUNCH(k_mf_transpose_dofs, grid_for(nc * nloc, TB), TB, 0, c->stream, nc, nloc,
                 s->cell_dofs.p, s->mf_cdT.p);
    const int64_t n = nc * nloc + nbf * nfd;
    SUCEM_CHECK(n < (int64_t(1) << 31), SUCEM_ERR_UNSUPPORTED, "too many element products");
    DBuf<uint32_t> keys(n), keys_out(n), vin(n), slot(n);
    SUCEM_LAUNCH(k_mf_keys, grid_for(n, TB), TB, 0, c->stream, nc, nloc, nbf, nfd, s->cell_dofs.p,
                 s->bf_cell.p, s->bf_local.p, s->ref_face_dofs.p, keys.p);
    SUCEM_LAUNCH(k_mf_iota, grid_for(n, TB), TB, 0, c->stream, n, vin.p);
    const unsigned bits = (unsigned)key_bits(s->ndofs);
    size_t bytes = 0;
    SUCEM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys.p, keys_out.p, vin.p, slot.p,
                                        (size_t)n, 0u, bits, c->stream));
    DBuf<char> tmp(bytes);
    SUCEM_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, keys.p, keys_out.p, vin.p, slot.p,
                                        (size_t)n, 0u, bits, c->stream));
    s->mf_dptr.alloc(s->ndofs + 1);
    SUCEM_LAUNCH(k_mf_heads, grid_for(n, TB), TB, 0, c->stream, n, s->ndofs, keys_out.p,
                 s->mf_dptr.p);
    s->mf_posT.alloc((size_t)nc * nloc);
    s->mf_posF.alloc((size_t)nbf * nfd);
    SUCEM_LAUNCH(k_mf_inverse, grid_for(n, TB), TB, 0, c->stream, n, nc, nloc, slot.p,
                 s->mf_posT.p, s->mf_posF.p);
    c->sync();   // temporaries are released on return
  }
  if (s->order == 2) mfma_fragments(c, s);
  sys->mf_coef64.alloc((size_t)12 * nc);
  sys->mf_coef32.alloc((size_t)12 * nc);
  SUCEM_LAUNCH((k_mf_coefs<double>), grid_for(nc, TB), TB, 0, c->stream, nc, m->xyz.p, m->cells.p,
               m->eps_r.p, m->inv_mu_r.p, sys->mf_coef64.p);
  SUCEM_LAUNCH((k_mf_coefs<float>), grid_for(nc, TB), TB, 0, c->stream, nc, m->xyz.p, m->cells.p,
               m->eps_r.p, m->inv_mu_r.p, sys->mf_coef32.p);
  sys->mf_face64.alloc((size_t)6 * nbf);
  sys->mf_face32.alloc((size_t)6 * nbf);
  if (nbf) {
    DBuf<double> H((size_t)6 * nbf);
    face_coefs(c, s, H.p);
    SUCEM_LAUNCH(k_mf_soa6, grid_for(nbf, TB), TB, 0, c->stream, nbf, H.p, sys->mf_face64.p,
                 sys->mf_face32.p);
    c->sync();
  }
  sys->mf_yl.alloc((size_t)(nc * nloc + nbf * nfd));
  c->sync();
  sys->mf_ready = true;
}

template <typename R>
void mf_apply(Context* c, const System* sys, Coef k, const uint8_t* mask, const void* x, void* y,
              const void* w, int epi, const RedSite& rs, const SpmvCfg& f) {
  using C = typename C2<R>::T;
  const Space* s = sys->space;
  const int64_t nc = s->mesh->nc, nbf = s->nbf;
  const R* coef;
  const R* hf;
  if constexpr (sizeof(R) == 4) {
    coef = sys->mf_coef32.p; hf = sys->mf_face32.p;
  } else {
    coef = sys->mf_coef64.p; hf = sys->mf_face64.p;
  }
  C* yl = (C*)sys->mf_yl.p;
  const int Gc = (int)((nc + 8 * TB - 1) / (8 * TB) * 8);   // TB cells per block, multiple of 8
  if (s->order == 2 && f.U == 3) {
    // element products on the matrix cores (FP32 only; 32 cells per block step)
    if constexpr (sizeof(R) == 4) {
      SUCEM_CHECK(s->mf_afrag.p, SUCEM_ERR_INVALID, "MFMA element tables not built");
      const int64_t ngroup = (nc + kMfmaCells - 1) / kMfmaCells;   // 32 cells per block step
      const int Gm = (int)std::max<int64_t>(1, std::min<int64_t>(ngroup,
                                                                   (int64_t)c->num_cus * 8));
      SUCEM_LAUNCH(k_mf_cells_mfma, Gm, TB, 0, c->stream, nc, coef, s->mf_cdT.p, s->mf_posT.p,
                   (float)k.s, (float)k.t, (const half8_t*)s->mf_afrag.p, (const float2*)x,
                   (float2*)yl);
    } else {
      SUCEM_CHECK(false, SUCEM_ERR_UNSUPPORTED, "the MFMA element products are FP32 only");
    }
  } else if (s->order == 2 && f.U == 2)
    SUCEM_LAUNCH((k_mf_cells<R, 2, true>), Gc, TB, 0, c->stream, nc, Gc, coef, s->mf_cdT.p,
                 s->mf_posT.p, (R)k.s, (R)k.t, (const C*)x, yl);
  else if (s->order == 2)
    SUCEM_LAUNCH((k_mf_cells<R, 2, false>), Gc, TB, 0, c->stream, nc, Gc, coef, s->mf_cdT.p,
                 s->mf_posT.p, (R)k.s, (R)k.t, (const C*)x, yl);
  else
    SUCEM_LAUNCH((k_mf_cells<R, 1, false>), Gc, TB, 0, c->stream, nc, Gc, coef, s->mf_cdT.p,
                 s->mf_posT.p, (R)k.s, (R)k.t, (const C*)x, yl);
  if (nbf) {
    // the ABC term only where S0 was assembled; other systems keep zero face products
    const R c0 = (sys->which & SUCEM_MAT_S0) ? (R)k.s0 : R(0);
    if (s->nfd == 8)
      SUCEM_LAUNCH((k_mf_faces<R, 8>), grid_for(nbf * 8, TB), TB, 0, c->stream, nbf, s->nloc, hf,
                   s->bf_cell.p, s->bf_local.p, s->cell_dofs.p, s->ref_face_dofs.p,
                   s->ref_face.p, c0, s->mf_posF.p, (const C*)x, yl);
    else
      SUCEM_LAUNCH((k_mf_faces<R, 3>), grid_for(nbf * 3, TB), TB, 0, c->stream, nbf, s->nloc, hf,
                   s->bf_cell.p, s->bf_local.p, s->cell_dofs.p, s->ref_face_dofs.p,
                   s->ref_face.p, c0, s->mf_posF.p, (const C*)x, yl);
  }
  const int n = (int)s->ndofs;
#define MF_ARGS n, f.G, s->mf_dptr.p, (const C*)yl, mask, (C*)y, (const C*)w, rs
  if (f.L == 0) {
    if (epi == 0) SUCEM_LAUNCH((k_mf_gather_lds<R, 0>), f.G, TB, 0, c->stream, MF_ARGS);
    else if (epi == 1) SUCEM_LAUNCH((k_mf_gather_lds<R, 1>), f.G, TB, 0, c->stream, MF_ARGS);
    else SUCEM_LAUNCH((k_mf_gather_lds<R, 2>), f.G, TB, 0, c->stream, MF_ARGS);
  } else if (f.L == 4) {
    if (epi == 0) SUCEM_LAUNCH((k_mf_gather<R, 0, 4>), f.G, TB, 0, c->stream, MF_ARGS);
    else if (epi == 1) SUCEM_LAUNCH((k_mf_gather<R, 1, 4>), f.G, TB, 0, c->stream, MF_ARGS);
    else SUCEM_LAUNCH((k_mf_gather<R, 2, 4>), f.G, TB, 0, c->stream, MF_ARGS);
  } else {
    if (epi == 0) SUCEM_LAUNCH((k_mf_gather<R, 0, 1>), f.G, TB, 0, c->stream, MF_ARGS);
    else if (epi == 1) SUCEM_LAUNCH((k_mf_gather<R, 1, 1>), f.G, TB, 0, c->stream, MF_ARGS);
    else SUCEM_LAUNCH((k_mf_gather<R, 2, 1>), f.G, TB, 0, c->stream, MF_ARGS);
  }
#undef MF_ARGS
}
template void mf_apply<float>(Context*, const System*, Coef, const uint8_t*, const void*, void*,
                              const void*, int, const RedSite&, const SpmvCfg&);
template void mf_apply<double>(Context*, const System*, Coef, const uint8_t*, const void*, void*,
                               const void*, int, const RedSite&, const SpmvCfg&);

}  // namespace sucem
