// Device helpers shared by the operator, Krylov and eigen translation units.
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

template <typename C>
__device__ __forceinline__ C from_d(double2 v);
template <> __device__ __forceinline__ float2 from_d<float2>(double2 v) {
  return make_float2((float)v.x, (float)v.y);
}
template <> __device__ __forceinline__ double2 from_d<double2>(double2 v) { return v; }

// a / b, or 0 when b == 0 (an exactly converged recursion then stops moving)
__device__ __forceinline__ double2 cdiv0(double2 a, double2 b) {
  double den = b.x * b.x + b.y * b.y;
  if (den == 0.0) return make_double2(0.0, 0.0);
  return make_double2((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}
__device__ __forceinline__ double2 cmuld(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Paired recursions: two real systems packed in one complex vector (system 0 in the real
// parts, system 1 in the imaginary parts) run as two independent real recursions over the
// same operator and preconditioner passes (linear with real coefficients, so the two parts
// never mix).  Their products are taken per part, (sum a.x b.x, sum a.y b.y), and their
// scalars (a, b) scale the parts separately.
template <typename T>
__device__ __forceinline__ T pmul(T a, T b) {
  return mk(a.x * b.x, a.y * b.y);
}
// per part a / b, 0 where b == 0
__device__ __forceinline__ double2 pdiv0(double2 a, double2 b) {
  return make_double2(b.x != 0.0 ? a.x / b.x : 0.0, b.y != 0.0 ? a.y / b.y : 0.0);
}
// a.b into d: the complex bilinear product, or per part when paired
__device__ __forceinline__ void dot_acc(double2& d, double2 a, double2 b, bool paired) {
  if (paired) {
    d.x += a.x * b.x;
    d.y += a.y * b.y;
  } else {
    d.x += a.x * b.x - a.y * b.y;
    d.y += a.x * b.y + a.y * b.x;
  }
}
// |a|^2 into d.x, or per part (d.x, d.y) when paired
__device__ __forceinline__ void sq_acc(double2& d, double2 a, bool paired) {
  if (paired) {
    d.x += a.x * a.x;
    d.y += a.y * a.y;
  } else {
    d.x += a.x * a.x + a.y * a.y;
  }
}

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

// Logical block index that gives each XCD a contiguous share of the rows
// (blocks are dispatched round-robin over the 8 XCDs; G is a multiple of 8).
__device__ __forceinline__ int xcd_block(int G) {
  const int b = blockIdx.x;
  return (b & 7) * (G >> 3) + (b >> 3);
}

// Position of `target` in the ascending col[lo, hi) (one CSR row), or -1.  The midpoint
// is lo + (hi - lo) / 2: the row pointers of a pattern with more than 2^30 nonzeros (the
// 120 x 60 x 90-box order-2 cavity has 1,078,395,480) make lo + hi overflow 32 bits.
__device__ __forceinline__ int find_col(const int32_t* col, int lo, int hi, int target) {
  hi -= 1;
  while (lo <= hi) {
    const int mid = lo + ((hi - lo) >> 1);
    const int cm = col[mid];
    if (cm == target) return mid;
    if (cm < target) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

// ---- device scalars of the Krylov recursions -------------------------------------------
enum ScalarSlot {
  S_RHO = 0,      // complex: current rho
  S_ALPHA = 2,    // complex
  S_BETA = 4,     // complex
  S_OMEGA = 6,    // complex (BiCGStab)
  S_RR = 8,       // |r|^2 of the recursion
  S_TRUE = 9,     // |b - A x|^2 (FP64)
  S_BNORM = 10,   // |b|^2 (paired recursions: the two systems' |b|^2 at 10 and 11)
  S_RHO_NEXT = 12,  // complex: rho for the next alpha (equal to S_RHO except between the
                    // fused p update and the next x/r update of patch-operator COCG)
  S_COUNT = 14
};

enum FinalizeMode {
  FIN_NONE = 0,
  FIN_COCG_ALPHA,   // s0 = p.q            -> alpha = rho / s0
  FIN_COCG_BETA,    // s0 = r.z, s1 = |r|^2 -> beta = s0/rho, rho = s0, hist
  FIN_RESID,        // s0 = |r|^2 (true), s1 = rho of the replaced residual
  FIN_NORM,         // s0 = |b|^2
  FIN_BICG_ALPHA,   // s0 = conj(rt).v      -> alpha = rho / s0
  FIN_BICG_OMEGA,   // s0 = conj(s).t, s1 = |t|^2 -> omega = conj(s0)/s1
  FIN_BICG_BETA,    // s0 = conj(rt).r, s1 = |r|^2 -> beta = (s0/rho)(alpha/omega), rho = s0
  FIN_SET_RHO,      // s0 = r.z, s1 = |r|^2     -> rho = s0 (start / restart)
  FIN_DEFER,        // partials only: the consuming kernel forms the scalars itself
};

// One reduction site: per-block partial sums and the scalar array they are finalised
// into (counter: unused, kept for layout).
struct RedSite {
  double2* part;      // 2*G entries
  unsigned* counter;  // one per site
  double* scal;       // S_COUNT doubles
  double* hist;       // residual history (nullable)
  int mode;
  int it;
  int paired = 0;     // two real systems packed as real and imaginary parts (see pmul)
};

// Paired recursions: the larger of the two systems' squared residual norms s.x, s.y, each
// relative to its |b|^2 (a system with b = 0 counts as solved).
__device__ __forceinline__ double prel2(const double* sc, double2 s) {
  const double b0 = sc[S_BNORM], b1 = sc[S_BNORM + 1];
  return fmax(b0 > 0.0 ? s.x / b0 : 0.0, b1 > 0.0 ? s.y / b1 : 0.0);
}

__device__ __forceinline__ void finalize(const RedSite& rs, double2 s0, double2 s1) {
  double* sc = rs.scal;
  const double2 rho = make_double2(sc[S_RHO], sc[S_RHO + 1]);
  double2 out;
  switch (rs.mode) {
    case FIN_COCG_ALPHA:
    case FIN_BICG_ALPHA:
      out = cdiv0(rho, s0);
      sc[S_ALPHA] = out.x; sc[S_ALPHA + 1] = out.y;
      break;
    case FIN_COCG_BETA:
      out = cdiv0(s0, rho);
      sc[S_BETA] = out.x; sc[S_BETA + 1] = out.y;
      sc[S_RHO] = s0.x; sc[S_RHO + 1] = s0.y;
      sc[S_RHO_NEXT] = s0.x; sc[S_RHO_NEXT + 1] = s0.y;
      sc[S_RR] = s1.x;
      if (rs.hist) rs.hist[rs.it] = sqrt(s1.x);
      break;
    case FIN_RESID:   // paired: S_TRUE holds the larger relative square (see prel2)
      sc[S_TRUE] = rs.paired ? prel2(sc, s0) : s0.x;
      sc[S_RHO] = s1.x; sc[S_RHO + 1] = s1.y;
      sc[S_RHO_NEXT] = s1.x; sc[S_RHO_NEXT + 1] = s1.y;
      break;
    case FIN_NORM:
      sc[S_BNORM] = s0.x;
      sc[S_BNORM + 1] = s0.y;   // paired: system 1 (0 otherwise)
      break;
    case FIN_SET_RHO:
      sc[S_RHO] = s0.x; sc[S_RHO + 1] = s0.y;
      sc[S_RHO_NEXT] = s0.x; sc[S_RHO_NEXT + 1] = s0.y;
      sc[S_RR] = s1.x;
      break;
    case FIN_BICG_OMEGA:
      out = s1.x == 0.0 ? make_double2(0.0, 0.0) : make_double2(s0.x / s1.x, -s0.y / s1.x);
      sc[S_OMEGA] = out.x; sc[S_OMEGA + 1] = out.y;
      break;
    case FIN_BICG_BETA: {
      const double2 al = make_double2(sc[S_ALPHA], sc[S_ALPHA + 1]);
      const double2 om = make_double2(sc[S_OMEGA], sc[S_OMEGA + 1]);
      out = cmuld(cdiv0(s0, rho), cdiv0(al, om));
      sc[S_BETA] = out.x; sc[S_BETA + 1] = out.y;
      sc[S_RHO] = s0.x; sc[S_RHO + 1] = s0.y;
      sc[S_RHO_NEXT] = s0.x; sc[S_RHO_NEXT + 1] = s0.y;
      sc[S_RR] = s1.x;
      if (rs.hist) rs.hist[rs.it] = sqrt(s1.x);
      break;
    }
    default:
      break;
  }
}

// Block sums of (v0, v1) -> partials part[lb] and part[G + lb]; k_finalize, launched
// after the producing kernel (finish_reduction), sums them in block order
// (deterministic) and finalises the scalars.  Ordering between the two kernels comes
// from the stream, so no block needs a device-scope fence (on a multi-XCD part each
// such fence writes back the XCD's L2).  Must be called by every thread of every block.
template <int TB>
__device__ __forceinline__ void reduce_finalize(const RedSite& rs, int G, double2 v0, double2 v1) {
  v0 = block_sum<TB>(v0);
  v1 = block_sum<TB>(v1);
  if (threadIdx.x == 0) {
    const int lb = xcd_block(G);
    rs.part[lb] = v0;
    rs.part[G + lb] = v1;
  }
}

__device__ __forceinline__ double2 scal2(const double* sc, int slot) {
  return make_double2(sc[slot], sc[slot + 1]);
}

}  // namespace sucem
