// Batched driven sweep: NB systems on one assembled system, solved together,
//
//   A_j x_j = b_j,   A_j = S - k0_j^2 T + i k0_j S0,   j < NB,
//
// (frequencies of a sweep and/or right-hand sides of one frequency; the reference re-forms
// and re-solves the LHS per frequency, sandbox/patch/patch_driver.py:106-129, and adds
// one source at a time, Sources/point_source.py:49-94).
//
// Vectors are interleaved [row][system].  One iteration of the NB COCG recursions
// (auxiliary-space preconditioner, as in the single-system solver) is:
//
//   k_sw_patch   q_j = A_j p_j for every j: one workgroup per patch of cells loads the
//                patch's maps, the cells' 12 geometric coefficients and LDS places once
//                and runs the gather / element product / segment-sum phases of the patch
//                operator for each system in turn (the per-system coefficients ct_j, c0_j
//                scale the same registers), plus the block partials of p_j.q_j;
//   k_sw_psum    rows of the DOFs shared by several patches (sums of partial sums);
//   k_sw_alpha   one block: alpha_j = rho_j / (p_j.q_j) (0 for a finished system);
//   k_sw_xr      x_j += alpha_j p_j, r_j -= alpha_j q_j;
//   k_sw_aux_t   vertex (and order-2 edge-node) rows of G^T r_j and Pi^T r_j, lanes =
//                (system, incidence stride), so a row's gathers read whole [row][NB] lines;
//   k_sw_aux_z   z_j = D_j^-1 r_j + theta G W_j G^T r_j + Pi V_j Pi^T r_j, partials of
//                r_j.z_j and |r_j|^2;
//   k_sw_beta    one block: beta_j, rho_j, |r_j| into a ring of the check interval;
//   k_sw_pz      p_j = z_j + beta_j p_j.
//
// The check interval's iterations (fixed launches, device-resident scalars) are captured
// once per solve in a hipGraph on a library stream and replayed; between replays the host
// reads the ring and, per system, does the FP64 residual replacement (FP32 recursions:
// whenever a system's recursion has dropped by ru_delta, or its cycle stagnates; FP64: at
// convergence), restarts the systems it replaced and retires the converged ones (their
// alpha and beta are 0 from then on, so x_j and r_j stay as checked).  Every sum has a
// fixed order: results are deterministic and do not depend on which systems share a batch.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "internal.hpp"
#include "kernels.hpp"
#include "mf_element.hpp"
#include "ops.hpp"

namespace sucem {
namespace {

constexpr int TB = 256;
constexpr int UBS = 8;       // patch DOFs per thread per gather / segment-sum batch
constexpr int SW = 8;        // per-system scalars: rho (2), alpha (2), beta (2), spare (2)
constexpr int SW_RHO = 0, SW_ALPHA = 2, SW_BETA = 4;

struct Rows { int start, end; };
__device__ __forceinline__ Rows sw_rows(int n, int G) {
  const int lb = xcd_block(G);
  const int chunk = (n + G - 1) / G;
  Rows r;
  r.start = lb * chunk;
  r.end = min(n, r.start + chunk);
  return r;
}

// Per-system block sums: every thread holds a value of system tid % NB; thread t < NB
// returns the block's sum for system t (fixed tree).  Contains barriers.
template <int NB>
__device__ __forceinline__ double2 block_sum_nb(double2 v) {
  __shared__ double2 red[TB];
  const int t = threadIdx.x;
  __syncthreads();
  red[t] = v;
  __syncthreads();
  for (int step = TB / 2; step >= NB; step >>= 1) {
    if (t < step) {
      red[t].x += red[t + step].x;
      red[t].y += red[t + step].y;
    }
    __syncthreads();
  }
  return red[t < NB ? t : 0];
}

// ---- operator -------------------------------------------------------------------
// Block sums of NBL values at once (two barriers for all of them); every thread gets them.
template <int CB, int NBL>
__device__ __forceinline__ void block_sum_multi(double2 (&v)[NBL]) {
  __shared__ double2 red[CB / 64][NBL];
#pragma unroll
  for (int jj = 0; jj < NBL; ++jj)
    for (int o = 32; o > 0; o >>= 1) {
      v[jj].x += __shfl_xor(v[jj].x, o);
      v[jj].y += __shfl_xor(v[jj].y, o);
    }
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int jj = 0; jj < NBL; ++jj) red[w][jj] = v[jj];
  __syncthreads();
#pragma unroll
  for (int jj = 0; jj < NBL; ++jj) {
    double2 t = red[0][jj];
    for (int i = 1; i < CB / 64; ++i) {
      t.x += red[i][jj].x;
      t.y += red[i][jj].y;
    }
    v[jj] = t;
  }
}

// One workgroup per patch; the systems in groups of NBL: each phase (gather of x, element
// products, segment sums) handles a whole group, with the patch's LDS copy laid out
// [slot][NBL].  The cell's coefficients and LDS places are loaded once for all systems.
// kco = {ct_0..ct_NB-1, c0_0..c0_NB-1} (cs = 1).  part: partial sums [slot][NB]; dp: with
// EPI, block partials of p_j.q_j at dp[j Gp + lb].
template <typename R, int ORDER, int NFD, int NB, int NBL, int CB, bool EPI>
__global__ __launch_bounds__(CB) void k_sw_patch(
    int64_t np, int Gp, int64_t nc, int64_t nbf, int lds_main, const R* __restrict__ coef,
    const uint32_t* __restrict__ lp, const int32_t* __restrict__ uptr,
    const int32_t* __restrict__ pdof, const uint16_t* __restrict__ pseg,
    const int32_t* __restrict__ pout, const int32_t* __restrict__ fptr,
    const int32_t* __restrict__ fmeta, const uint16_t* __restrict__ fpos,
    const R* __restrict__ hf, const int32_t* __restrict__ face_dofs,
    const double* __restrict__ ref_face, const double* __restrict__ kco,
    const uint8_t* __restrict__ mask, const typename C2<R>::T* __restrict__ x,
    typename C2<R>::T* __restrict__ y, typename C2<R>::T* __restrict__ part,
    double2* __restrict__ dp) {
  using C = typename C2<R>::T;
  static_assert(NB % NBL == 0, "groups of NBL systems");
  constexpr int NLOC = MfTab<ORDER>::NLOC;
  constexpr int KK = NFD * NFD;
  extern __shared__ double2 smem_[];
  C* sm = reinterpret_cast<C*>(smem_);
  const int lb = xcd_block(Gp);
  const int64_t p = lb;
  const bool live = p < np;   // uniform over the workgroup
  int u0 = 0, nu = 0, f0 = 0, nfr = 0;
  const int64_t r = p * CB + threadIdx.x;
  bool has = false;
  if (live) {
    u0 = uptr[p];
    nu = uptr[p + 1] - u0;
    f0 = fptr[p];
    nfr = (fptr[p + 1] - f0) * NFD;
    has = r < nc;
  }
  R pa[12];
  uint32_t pli[NLOC];
  if (has) {
#pragma unroll
    for (int k = 0; k < 12; ++k) pa[k] = coef[k * nc + r];
#pragma unroll
    for (int i = 0; i < NLOC; ++i) pli[i] = lp[(int64_t)i * nc + r];
  }
  C* stage = sm + (int64_t)lds_main * NBL;
  for (int jb = 0; jb < NB; jb += NBL) {
    double2 d1[NBL];
#pragma unroll
    for (int jj = 0; jj < NBL; ++jj) d1[jj] = make_double2(0.0, 0.0);
    if (live) {
      for (int j0 = 0; j0 < nu; j0 += UBS * CB) {
        int dv[UBS];
#pragma unroll
        for (int q = 0; q < UBS; ++q) {
          const int jj = j0 + q * CB + threadIdx.x;
          dv[q] = jj < nu ? pdof[u0 + jj] : -1;
        }
#pragma unroll
        for (int q = 0; q < UBS; ++q) {
          const int u = j0 + q * CB + threadIdx.x;
          if (dv[q] >= 0) {
            C xv[NBL];
#pragma unroll
            for (int jj = 0; jj < NBL; ++jj) xv[jj] = x[(int64_t)dv[q] * NB + jb + jj];
#pragma unroll
            for (int jj = 0; jj < NBL; ++jj) sm[u * NBL + jj] = xv[jj];
          }
        }
      }
    }
    __syncthreads();
    C acc[NBL][NLOC];
    if (has) {
#pragma unroll
      for (int jj = 0; jj < NBL; ++jj) {
        const R ct = (R)kco[jb + jj];
        R a[12];
#pragma unroll
        for (int k = 0; k < 12; ++k)
          a[k] = pa[k] * (k < 6 ? ct : R(1)) * R(1.0 / MfTab<ORDER>::den);
        C xl[NLOC];
#pragma unroll
        for (int i = 0; i < NLOC; ++i) xl[i] = sm[(pli[i] & 0xffffu) * NBL + jj];
#pragma unroll
        for (int i = 0; i < NLOC; ++i) acc[jj][i] = mk(R(0), R(0));
        element_product<R, ORDER>(a, xl, acc[jj],
                                  std::make_integer_sequence<int, MfTab<ORDER>::NP>{});
      }
    }
    for (int id = threadIdx.x; id < nfr; id += CB) {
      const int64_t fp = f0 + id / NFD;
      const int aa = id % NFD;
      const int mt = fmeta[fp];
      const int64_t rk = mt >> 2;
      const int lf = mt & 3;
      R hk[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) hk[k] = hf[k * nbf + fp];
      const double* rf = ref_face + lf * 6 * KK + aa * NFD;
      C v[NBL];
#pragma unroll
      for (int jj = 0; jj < NBL; ++jj) v[jj] = mk(R(0), R(0));
#pragma unroll
      for (int b = 0; b < NFD; ++b) {
        R m = R(0);
#pragma unroll
        for (int k = 0; k < 6; ++k) m += hk[k] * (R)rf[k * KK + b];
        const uint32_t li = lp[(int64_t)face_dofs[lf * NFD + b] * nc + rk] & 0xffffu;
#pragma unroll
        for (int jj = 0; jj < NBL; ++jj) {
          const C xb = sm[li * NBL + jj];
          v[jj].x += m * xb.x;
          v[jj].y += m * xb.y;
        }
      }
#pragma unroll
      for (int jj = 0; jj < NBL; ++jj) {
        const R c0 = (R)kco[NB + jb + jj];
        stage[id * NBL + jj] = mk(-c0 * v[jj].y, c0 * v[jj].x);
      }
    }
    __syncthreads();   // the x copy is no longer read: the products overwrite it
    if (has) {
#pragma unroll
      for (int i = 0; i < NLOC; ++i)
#pragma unroll
        for (int jj = 0; jj < NBL; ++jj) sm[(pli[i] >> 16) * NBL + jj] = acc[jj][i];
    }
    for (int id = threadIdx.x; id < nfr; id += CB) {
      const int t = fpos[(int64_t)f0 * NFD + id];
#pragma unroll
      for (int jj = 0; jj < NBL; ++jj) sm[t * NBL + jj] = stage[id * NBL + jj];
    }
    __syncthreads();
    if (live) {
      for (int j0 = 0; j0 < nu; j0 += UBS * CB) {
        int lo[UBS], hi[UBS], out[UBS], dv[UBS];
#pragma unroll
        for (int q = 0; q < UBS; ++q) {
          const int jj = j0 + q * CB + threadIdx.x;
          const bool in = jj < nu;
          lo[q] = (in && jj) ? (int)pseg[u0 + jj - 1] : 0;
          hi[q] = in ? (int)pseg[u0 + jj] : 0;
          out[q] = in ? pout[u0 + jj] : 0;
          dv[q] = in ? pdof[u0 + jj] : 0;
        }
#pragma unroll
        for (int q = 0; q < UBS; ++q) {
          if (hi[q] == 0) continue;   // not an entry of this patch (segments are never empty)
          C v[NBL];
#pragma unroll
          for (int jj = 0; jj < NBL; ++jj) v[jj] = mk(R(0), R(0));
          for (int t = lo[q]; t < hi[q]; ++t) {
#pragma unroll
            for (int jj = 0; jj < NBL; ++jj) {
              const C s = sm[t * NBL + jj];
              v[jj].x += s.x;
              v[jj].y += s.y;
            }
          }
          const bool masked = out[q] < 0 && mask && mask[dv[q]];
#pragma unroll
          for (int jj = 0; jj < NBL; ++jj) {
            C w = v[jj];
            if (out[q] >= 0) {
              part[(int64_t)out[q] * NB + jb + jj] = w;
            } else {
              if (masked) w = mk(R(0), R(0));
              y[(int64_t)dv[q] * NB + jb + jj] = w;
            }
            if constexpr (EPI) {
              // p.Ap over the finished rows and the partial sums (linear in q; p is zero
              // on masked rows)
              const double2 xd = to_d(x[(int64_t)dv[q] * NB + jb + jj]), yd = to_d(w);
              d1[jj].x += xd.x * yd.x - xd.y * yd.y;
              d1[jj].y += xd.x * yd.y + xd.y * yd.x;
            }
          }
        }
      }
    }
    if constexpr (EPI) {
      block_sum_multi<CB, NBL>(d1);   // its barriers also separate this group's reads of the
      if (threadIdx.x == 0)            // products from the next group's gather
#pragma unroll
        for (int jj = 0; jj < NBL; ++jj) dp[(int64_t)(jb + jj) * Gp + lb] = d1[jj];
    } else {
      __syncthreads();
    }
  }
}

// Rows of the DOFs shared by several patches: y = sum of their partial sums, in patch order.
template <typename R, int NB>
__global__ __launch_bounds__(TB) void k_sw_psum(int npd, int G, const int32_t* __restrict__ qptr,
                                                const int32_t* __restrict__ qdof,
                                                const typename C2<R>::T* __restrict__ part,
                                                const uint8_t* __restrict__ mask,
                                                typename C2<R>::T* __restrict__ y) {
  using C = typename C2<R>::T;
  const Rows rr = sw_rows(npd, G);
  for (int64_t e = (int64_t)rr.start * NB + threadIdx.x; e < (int64_t)rr.end * NB; e += TB) {
    const int i = (int)(e / NB), j = (int)(e % NB);
    C v = mk(R(0), R(0));
    const int q1 = qptr[i + 1];
    for (int q = qptr[i]; q < q1; ++q) {
      const C t = part[(int64_t)q * NB + j];
      v.x += t.x;
      v.y += t.y;
    }
    const int d = qdof[i];
    if (mask && mask[d]) v = mk(R(0), R(0));
    y[(int64_t)d * NB + j] = v;
  }
}

// ---- scalars: one block, one wavefront per system ------------------------------------
// Lane l of wavefront j sums system j's block partials l, l + 64, ... (four independent
// running sums, combined in a fixed order), then the wavefront reduces: the order of every
// sum is fixed, and the systems are summed in parallel.
__device__ __forceinline__ double2 wave_sum_partials(const double2* __restrict__ a, int G) {
  const int lane = threadIdx.x & 63;
  double2 s[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) s[k] = make_double2(0.0, 0.0);
  int b = lane;
  for (; b + 192 < G; b += 256) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double2 t = a[b + 64 * k];
      s[k].x += t.x;
      s[k].y += t.y;
    }
  }
  for (int k = 0; b < G; b += 64, ++k) {
    const double2 t = a[b];
    s[k & 3].x += t.x;
    s[k & 3].y += t.y;
  }
  double2 v = make_double2((s[0].x + s[1].x) + (s[2].x + s[3].x),
                           (s[0].y + s[1].y) + (s[2].y + s[3].y));
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o);
    v.y += __shfl_xor(v.y, o);
  }
  return v;
}

template <int NB>
__global__ __launch_bounds__(64 * NB) void k_sw_alpha(int Gp, const double2* __restrict__ dp,
                                                      const int* __restrict__ act,
                                                      double* __restrict__ sc) {
  const int j = threadIdx.x >> 6;
  const double2 s = wave_sum_partials(dp + (int64_t)j * Gp, Gp);
  if ((threadIdx.x & 63) == 0) {
    const double2 al = act[j] ? cdiv0(scal2(sc + j * SW, SW_RHO), s) : make_double2(0, 0);
    sc[j * SW + SW_ALPHA] = al.x;
    sc[j * SW + SW_ALPHA + 1] = al.y;
  }
}

// mode 0 (iteration): beta_j = (r.z)_j / rho_j, rho_j = (r.z)_j, ring[k NB + j] = |r_j| for
// the running systems (beta 0 for the others); mode 1 (restart): rho_j = (r.z)_j for the
// systems with sel[j].
template <int NB>
__global__ __launch_bounds__(64 * NB) void k_sw_beta(int G, const double2* __restrict__ dp,
                                                     const int* __restrict__ act,
                                                     const int* __restrict__ sel, double* sc,
                                                     double* __restrict__ ring, int k, int mode) {
  const int j = threadIdx.x >> 6;
  const double2 s0 = wave_sum_partials(dp + (int64_t)j * G, G);
  const double2 s1 = wave_sum_partials(dp + (int64_t)(NB + j) * G, G);
  if ((threadIdx.x & 63) == 0) {
    double* s = sc + j * SW;
    if (mode == 0) {
      const bool on = act[j] != 0;
      const double2 be = on ? cdiv0(s0, scal2(s, SW_RHO)) : make_double2(0, 0);
      s[SW_BETA] = be.x;
      s[SW_BETA + 1] = be.y;
      if (on) {
        s[SW_RHO] = s0.x;
        s[SW_RHO + 1] = s0.y;
      }
      ring[k * NB + j] = sqrt(s1.x);
    } else if (sel[j]) {
      s[SW_RHO] = s0.x;
      s[SW_RHO + 1] = s0.y;
    }
  }
}

// sums[j] = sum of the block partials dp[j G + b] (.x), for norms
template <int NB>
__global__ __launch_bounds__(64 * NB) void k_sw_sum(int G, const double2* __restrict__ dp,
                                                    double* __restrict__ sums) {
  const int j = threadIdx.x >> 6;
  const double2 s = wave_sum_partials(dp + (int64_t)j * G, G);
  if ((threadIdx.x & 63) == 0) sums[j] = s.x;
}

// ---- vector kernels (element e = row NB + j; TB is a multiple of NB, so a thread keeps
// its system) ---------------------------------------------------------------------------
template <typename R, int NB>
__global__ __launch_bounds__(TB) void k_sw_xr(int n, int G, const double* __restrict__ sc,
                                              typename C2<R>::T* __restrict__ x,
                                              typename C2<R>::T* __restrict__ r,
                                              const typename C2<R>::T* __restrict__ p,
                                              const typename C2<R>::T* __restrict__ q) {
  using C = typename C2<R>::T;
  const Rows rr = sw_rows(n, G);
  const int j = threadIdx.x % NB;
  const C a = from_d<C>(scal2(sc + j * SW, SW_ALPHA));
  for (int64_t e = (int64_t)rr.start * NB + threadIdx.x; e < (int64_t)rr.end * NB; e += TB) {
    const C ap = cmul(a, p[e]), aq = cmul(a, q[e]);
    C xi = x[e], ri = r[e];
    xi.x += ap.x; xi.y += ap.y;
    ri.x -= aq.x; ri.y -= aq.y;
    x[e] = xi;
    r[e] = ri;
  }
}

template <typename R, int NB>
__global__ __launch_bounds__(TB) void k_sw_pz(int n, int G, const double* __restrict__ sc,
                                              const typename C2<R>::T* __restrict__ z,
                                              typename C2<R>::T* __restrict__ p) {
  using C = typename C2<R>::T;
  const Rows rr = sw_rows(n, G);
  const int j = threadIdx.x % NB;
  const C b = from_d<C>(scal2(sc + j * SW, SW_BETA));
  for (int64_t e = (int64_t)rr.start * NB + threadIdx.x; e < (int64_t)rr.end * NB; e += TB) {
    const C bp = cmul(b, p[e]), zi = z[e];
    p[e] = mk(zi.x + bp.x, zi.y + bp.y);
  }
}

// p_j = z_j for the restarted systems
template <typename R, int NB>
__global__ __launch_bounds__(TB) void k_sw_restart_p(int n, int G, const int* __restrict__ sel,
                                                     const typename C2<R>::T* __restrict__ z,
                                                     typename C2<R>::T* __restrict__ p) {
  const Rows rr = sw_rows(n, G);
  if (!sel[threadIdx.x % NB]) return;
  for (int64_t e = (int64_t)rr.start * NB + threadIdx.x; e < (int64_t)rr.end * NB; e += TB)
    p[e] = z[e];
}

// FP32 refinement, selected systems: x64 += sign * xlo (the cycle's correction, or its
// removal), then xlo = 0 if clear
template <typename R, int NB>
__global__ __launch_bounds__(TB) void k_sw_accum(int n, int G, const int* __restrict__ sel,
                                                 typename C2<R>::T* __restrict__ xlo,
                                                 double2* __restrict__ x64, double sign,
                                                 int clear) {
  const Rows rr = sw_rows(n, G);
  if (!sel[threadIdx.x % NB]) return;
  for (int64_t e = (int64_t)rr.start * NB + threadIdx.x; e < (int64_t)rr.end * NB; e += TB) {
    const double2 a = to_d(xlo[e]);
    x64[e].x += sign * a.x;
    x64[e].y += sign * a.y;
    if (clear) xlo[e] = mk(R(0), R(0));
  }
}

// selected systems: r = b - ax (masked rows 0) in the working precision; partials of the
// FP64 |r|^2 (every system's partials are written; unselected ones contribute 0)
template <typename R, int NB>
__global__ __launch_bounds__(TB) void k_sw_resid(int n, int G, const int* __restrict__ sel,
                                                 const double2* __restrict__ b,
                                                 const double2* __restrict__ ax,
                                                 const uint8_t* __restrict__ mask,
                                                 typename C2<R>::T* __restrict__ r,
                                                 double2* __restrict__ dp) {
  using C = typename C2<R>::T;
  const Rows rr = sw_rows(n, G);
  const int j = threadIdx.x % NB;
  const bool on = sel[j] != 0;
  double2 s = make_double2(0.0, 0.0);
  if (on) {
    for (int64_t e = (int64_t)rr.start * NB + threadIdx.x; e < (int64_t)rr.end * NB; e += TB) {
      double2 v = make_double2(b[e].x - ax[e].x, b[e].y - ax[e].y);
      if (mask && mask[e / NB]) v = make_double2(0.0, 0.0);
      s.x += v.x * v.x + v.y * v.y;
      r[e] = from_d<C>(v);
    }
  }
  s = block_sum_nb<NB>(s);
  if (threadIdx.x < NB) dp[(int64_t)threadIdx.x * G + xcd_block(G)] = s;
}

// |b_j|^2 partials (masked rows excluded)
template <int NB>
__global__ __launch_bounds__(TB) void k_sw_norm(int n, int G, const double2* __restrict__ b,
                                                const uint8_t* __restrict__ mask,
                                                double2* __restrict__ dp) {
  const Rows rr = sw_rows(n, G);
  double2 s = make_double2(0.0, 0.0);
  for (int64_t e = (int64_t)rr.start * NB + threadIdx.x; e < (int64_t)rr.end * NB; e += TB)
    if (!(mask && mask[e / NB])) s.x += b[e].x * b[e].x + b[e].y * b[e].y;
  s = block_sum_nb<NB>(s);
  if (threadIdx.x < NB) dp[(int64_t)threadIdx.x * G + xcd_block(G)] = s;
}

// out[row NB + j] = in_j[row] (system-major, nb systems; systems j >= nb get zeros)
template <typename C>
__global__ void k_sw_interleave(int64_t n, int NB, int nb, const C* __restrict__ in,
                                C* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * NB;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / NB;
    const int j = (int)(e % NB);
    out[e] = j < nb ? in[(int64_t)j * n + i] : from_d<C>(make_double2(0.0, 0.0));
  }
}

// one system's vector (n entries) into column j of an interleaved array
template <typename C>
__global__ void k_sw_put(int64_t n, int NB, int j, const C* __restrict__ in, C* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i * NB + j] = in[i];
}

__global__ void k_sw_deinterleave(int64_t n, int NB, int nb, const double2* __restrict__ in,
                                  double2* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * NB;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / NB;
    const int j = (int)(e % NB);
    if (j < nb) out[(int64_t)j * n + i] = in[e];
  }
}

// zero n 32-bit words
__global__ void k_sw_zero(int64_t n, uint32_t* __restrict__ a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0u;
}

template <typename R>
__global__ void k_sw_to_lo(int64_t n, const double2* __restrict__ in,
                           typename C2<R>::T* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = from_d<typename C2<R>::T>(in[i]);
}

// ---- auxiliary-space preconditioner on interleaved vectors -------------------------
// The same nodal formulas as the single-system kernels (hiptmair.hip): G and Pi are
// applied from entity ids, adjacency lists and coordinates, nothing stored.
template <typename R, int NB, int LR, int ORDER>
__device__ __forceinline__ void sw_vertex_rows(int64_t blk, int nv, int nE, R inv_den,
                                               const int32_t* __restrict__ ve_ptr,
                                               const int32_t* __restrict__ ve_ent,
                                               const int32_t* __restrict__ vf_ptr,
                                               const int32_t* __restrict__ vf_ent,
                                               const int32_t* __restrict__ edges,
                                               const int32_t* __restrict__ faces,
                                               const R* __restrict__ xyz,
                                               const typename C2<R>::T* __restrict__ r,
                                               const typename C2<R>::T* __restrict__ winv,
                                               const typename C2<R>::T* __restrict__ vinv,
                                               typename C2<R>::T* __restrict__ w,
                                               typename C2<R>::T* __restrict__ u) {
  using C = typename C2<R>::T;
  constexpr int L = NB * LR;
  const int lane = threadIdx.x % L;
  const int j = lane % NB, lr = lane / NB;
  const int64_t v = (blk * TB + threadIdx.x) / L;
  C g = mk(R(0), R(0));
  C acc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) acc[d] = mk(R(0), R(0));
  if (v < nv) {
    R xv[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) xv[d] = xyz[3 * v + d];
    for (int q = ve_ptr[v] + lr; q < ve_ptr[v + 1]; q += LR) {
      const int ent = ve_ent[q];
      const int e = ent >> 1;
      const bool second = ent & 1;
      C sr;
      if constexpr (ORDER == 2) {
        const C r0 = r[(int64_t)(2 * e) * NB + j], r1 = r[(int64_t)(2 * e + 1) * NB + j];
        const R g0 = second ? R(1) : R(-5), g1 = second ? R(5) : R(-1);
        g.x += g0 * r0.x + g1 * r1.x;
        g.y += g0 * r0.y + g1 * r1.y;
        const R l0 = second ? R(1.0 / 3.0) : R(2.0 / 3.0), l1 = R(1) - l0;
        sr = mk(l0 * r0.x + l1 * r1.x, l0 * r0.y + l1 * r1.y);
      } else {
        const C r0 = r[(int64_t)e * NB + j];
        g.x += second ? r0.x : -r0.x;
        g.y += second ? r0.y : -r0.y;
        sr = mk(R(0.5) * r0.x, R(0.5) * r0.y);
      }
      const int o = edges[2 * e + (second ? 0 : 1)];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const R xo = xyz[3 * o + d];
        const R t = second ? xv[d] - xo : xo - xv[d];
        acc[d].x += t * sr.x;
        acc[d].y += t * sr.y;
      }
    }
    if constexpr (ORDER == 2) {
      for (int q = vf_ptr[v] + lr; q < vf_ptr[v + 1]; q += LR) {
        const int ent = vf_ent[q];
        const int f = ent >> 2, pos = ent & 3;
        const C r0 = r[(int64_t)(2 * nE + 2 * f) * NB + j];
        const C r1 = r[(int64_t)(2 * nE + 2 * f + 1) * NB + j];
        if (pos == 0) { g.x -= r0.x + r1.x; g.y -= r0.y + r1.y; }
        else if (pos == 1) { g.x += r0.x; g.y += r0.y; }
        else { g.x += r1.x; g.y += r1.y; }
        const int o1 = faces[3 * f + (pos == 0 ? 1 : 0)], o2 = faces[3 * f + (pos == 2 ? 1 : 2)];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const R y1 = xyz[3 * o1 + d], y2 = xyz[3 * o2 + d];
          const R xp = pos == 0 ? xv[d] : y1;
          const R xq = pos == 0 ? y1 : (pos == 1 ? xv[d] : y2);
          const R xr = pos == 2 ? xv[d] : y2;
          const R t0 = (xq - xp) * R(1.0 / 3.0);
          const R t1 = (xr - xp) * R(1.0 / 3.0);
          acc[d].x += t0 * r0.x + t1 * r1.x;
          acc[d].y += t0 * r0.y + t1 * r1.y;
        }
      }
    }
  }
#pragma unroll
  for (int o = L / 2; o >= NB; o >>= 1) {
    g.x += __shfl_xor(g.x, o);
    g.y += __shfl_xor(g.y, o);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      acc[d].x += __shfl_xor(acc[d].x, o);
      acc[d].y += __shfl_xor(acc[d].y, o);
    }
  }
  if (lr == 0 && v < nv) {
    w[v * NB + j] = cmul(winv[v * NB + j], mk(inv_den * g.x, inv_den * g.y));
#pragma unroll
    for (int d = 0; d < 3; ++d)
      u[(3 * v + d) * NB + j] = cmul(vinv[(3 * v + d) * NB + j], acc[d]);
  }
}

template <typename R, int NB, int LR>
__device__ __forceinline__ void sw_edge_rows(int64_t blk, int nv, int nE, R inv_den,
                                             const int32_t* __restrict__ ef_ptr,
                                             const int32_t* __restrict__ ef_ent,
                                             const typename C2<R>::T* __restrict__ r,
                                             const typename C2<R>::T* __restrict__ winv,
                                             typename C2<R>::T* __restrict__ w) {
  using C = typename C2<R>::T;
  constexpr int L = NB * LR;
  const int lane = threadIdx.x % L;
  const int j = lane % NB, lr = lane / NB;
  const int64_t e = (blk * TB + threadIdx.x) / L;
  C g = mk(R(0), R(0));
  if (e < nE) {
    if (lr == 0) {
      const C r0 = r[(2 * e) * NB + j], r1 = r[(2 * e + 1) * NB + j];
      g = mk(R(4) * (r0.x - r1.x), R(4) * (r0.y - r1.y));
    }
    for (int q = ef_ptr[e] + lr; q < ef_ptr[e + 1]; q += LR) {
      const int ent = ef_ent[q];
      const int f = ent >> 2, pos = ent & 3;
      const C r0 = r[(int64_t)(2 * nE + 2 * f) * NB + j];
      const C r1 = r[(int64_t)(2 * nE + 2 * f + 1) * NB + j];
      if (pos == 0) { g.x -= R(4) * r1.x; g.y -= R(4) * r1.y; }
      else if (pos == 1) { g.x -= R(4) * r0.x; g.y -= R(4) * r0.y; }
      else { g.x += R(4) * (r0.x + r1.x); g.y += R(4) * (r0.y + r1.y); }
    }
  }
#pragma unroll
  for (int o = L / 2; o >= NB; o >>= 1) {
    g.x += __shfl_xor(g.x, o);
    g.y += __shfl_xor(g.y, o);
  }
  if (lr == 0 && e < nE)
    w[(nv + e) * NB + j] = cmul(winv[(nv + e) * NB + j], mk(inv_den * g.x, inv_den * g.y));
}

// blocks [0, Gv): vertex rows; [Gv, gridDim): order-2 edge-node rows (XCD-contiguous)
template <typename R, int NB, int LV, int LE, int ORDER>
__global__ __launch_bounds__(TB) void k_sw_aux_t(int Gv, int nv, int nE, R inv_den,
                                                 const int32_t* __restrict__ ve_ptr,
                                                 const int32_t* __restrict__ ve_ent,
                                                 const int32_t* __restrict__ vf_ptr,
                                                 const int32_t* __restrict__ vf_ent,
                                                 const int32_t* __restrict__ ef_ptr,
                                                 const int32_t* __restrict__ ef_ent,
                                                 const int32_t* __restrict__ edges,
                                                 const int32_t* __restrict__ faces,
                                                 const R* __restrict__ xyz,
                                                 const typename C2<R>::T* __restrict__ r,
                                                 const typename C2<R>::T* __restrict__ winv,
                                                 const typename C2<R>::T* __restrict__ vinv,
                                                 typename C2<R>::T* __restrict__ w,
                                                 typename C2<R>::T* __restrict__ u) {
  const int b = blockIdx.x;
  if (b < Gv) {
    const int lb = (b & 7) * (Gv >> 3) + (b >> 3);
    sw_vertex_rows<R, NB, LV, ORDER>(lb, nv, nE, inv_den, ve_ptr, ve_ent, vf_ptr, vf_ent, edges,
                                     faces, xyz, r, winv, vinv, w, u);
  } else if constexpr (ORDER == 2) {
    const int be = b - Gv, Ge = (int)gridDim.x - Gv;
    const int lb = (be & 7) * (Ge >> 3) + (be >> 3);
    sw_edge_rows<R, NB, LE>(lb, nv, nE, inv_den, ef_ptr, ef_ent, r, winv, w);
  }
}

// (G w_j)_i and (Pi u_j)_i
template <typename R, int NB, int ORDER>
__device__ __forceinline__ void sw_aux_row(int i, int j, int nv, int nE,
                                           const int32_t* __restrict__ edges,
                                           const int32_t* __restrict__ faces,
                                           const int32_t* __restrict__ face_edges,
                                           const R* __restrict__ xyz,
                                           const typename C2<R>::T* __restrict__ w,
                                           const typename C2<R>::T* __restrict__ u,
                                           typename C2<R>::T& gw, typename C2<R>::T& pu) {
  using C = typename C2<R>::T;
  gw = mk(R(0), R(0));
  pu = mk(R(0), R(0));
  const int ne_dofs = ORDER == 2 ? 2 * nE : nE;
  if (i < ne_dofs) {
    const int e = ORDER == 2 ? i >> 1 : i;
    const int k = ORDER == 2 ? (i & 1) : 0;
    const int a = edges[2 * e], b = edges[2 * e + 1];
    const C wa = w[(int64_t)a * NB + j], wb = w[(int64_t)b * NB + j];
    if constexpr (ORDER == 2) {
      const C we = w[(int64_t)(nv + e) * NB + j];
      const R ca = k ? R(-1) : R(-5), cb = k ? R(5) : R(1), ce = k ? R(-4) : R(4);
      gw = mk(ca * wa.x + cb * wb.x + ce * we.x, ca * wa.y + cb * wb.y + ce * we.y);
    } else {
      gw = mk(wb.x - wa.x, wb.y - wa.y);
    }
    const R lb = ORDER == 2 ? R(k + 1) * R(1.0 / 3.0) : R(0.5);
    const R la = R(1) - lb;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const R t = xyz[3 * b + d] - xyz[3 * a + d];
      const C ua = u[(int64_t)(3 * a + d) * NB + j], ub = u[(int64_t)(3 * b + d) * NB + j];
      pu.x += t * (la * ua.x + lb * ub.x);
      pu.y += t * (la * ua.y + lb * ub.y);
    }
  } else if constexpr (ORDER == 2) {
    const int f = (i - ne_dofs) >> 1;
    const int k = i & 1;
    const int p = faces[3 * f], q = faces[3 * f + 1], r = faces[3 * f + 2];
    const int o = k ? r : q;
    const C wp = w[(int64_t)p * NB + j], wo = w[(int64_t)o * NB + j];
    const C we = w[(int64_t)(nv + face_edges[3 * f + (k ? 0 : 1)]) * NB + j];
    const C wqr = w[(int64_t)(nv + face_edges[3 * f + 2]) * NB + j];
    gw = mk(wo.x - wp.x + R(4) * (wqr.x - we.x), wo.y - wp.y + R(4) * (wqr.y - we.y));
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const R t = (xyz[3 * o + d] - xyz[3 * p + d]) * R(1.0 / 3.0);
      const C up = u[(int64_t)(3 * p + d) * NB + j], uq = u[(int64_t)(3 * q + d) * NB + j];
      const C ur = u[(int64_t)(3 * r + d) * NB + j];
      pu.x += t * (up.x + uq.x + ur.x);
      pu.y += t * (up.y + uq.y + ur.y);
    }
  }
}

// z = dinv r + inv_den G w + Pi u (masked rows 0); partials of r.z (dp[j G + lb]) and
// |r|^2 (dp[(NB + j) G + lb]) per system
template <typename R, int NB, int ORDER>
__global__ __launch_bounds__(TB) void k_sw_aux_z(int n, int G, int nv, int nE, R inv_den,
                                                 const int32_t* __restrict__ edges,
                                                 const int32_t* __restrict__ faces,
                                                 const int32_t* __restrict__ face_edges,
                                                 const R* __restrict__ xyz,
                                                 const typename C2<R>::T* __restrict__ w,
                                                 const typename C2<R>::T* __restrict__ u,
                                                 const typename C2<R>::T* __restrict__ dinv,
                                                 const typename C2<R>::T* __restrict__ r,
                                                 typename C2<R>::T* __restrict__ z,
                                                 const uint8_t* __restrict__ mask,
                                                 double2* __restrict__ dp) {
  using C = typename C2<R>::T;
  const Rows rr = sw_rows(n, G);
  const int j = threadIdx.x % NB;
  double2 d1 = make_double2(0, 0), d2 = make_double2(0, 0);
  for (int64_t e = (int64_t)rr.start * NB + threadIdx.x; e < (int64_t)rr.end * NB; e += TB) {
    const int i = (int)(e / NB);
    const C ri = r[e];
    C gw, pu;
    sw_aux_row<R, NB, ORDER>(i, j, nv, nE, edges, faces, face_edges, xyz, w, u, gw, pu);
    const C di = cmul(dinv[e], ri);
    C zi = mk(di.x + inv_den * gw.x + pu.x, di.y + inv_den * gw.y + pu.y);
    if (mask && mask[i]) zi = mk(R(0), R(0));
    z[e] = zi;
    const double2 rd = to_d(ri), zd = to_d(zi);
    d1.x += rd.x * zd.x - rd.y * zd.y;
    d1.y += rd.x * zd.y + rd.y * zd.x;
    d2.x += rd.x * rd.x + rd.y * rd.y;
  }
  d1 = block_sum_nb<NB>(d1);
  if (threadIdx.x < NB) dp[(int64_t)threadIdx.x * G + xcd_block(G)] = d1;
  d2 = block_sum_nb<NB>(d2);
  if (threadIdx.x < NB) dp[(int64_t)(NB + threadIdx.x) * G + xcd_block(G)] = d2;
}

// ---- host driver --------------------------------------------------------------------
int patch_cb(int order, bool fp32) { return (order == 2 && !fp32) ? 192 : 256; }

// Everything one chunk of NB systems needs on the device.
template <typename R, int NB, int ORDER>
struct Chunk {
  using C = typename C2<R>::T;
  static constexpr int NFD = ORDER == 2 ? 8 : 3;
  // systems per phase of the working-precision patch kernel: order 1 takes four at a time
  // (its patch copy is small: ~60 KB of LDS for four FP32 systems); order 2 one
  static constexpr int NBL = (ORDER == 1 && NB >= 4) ? 4 : 1;
  static constexpr int LV = NB >= 8 ? 2 : (NB >= 4 ? 4 : 8 / NB);  // lanes = NB * LV <= 16
  static constexpr int LE = NB >= 2 ? 1 : 2;
  Context* c;
  System* sys;
  const Space* s;
  hipStream_t st;
  int n, G, Gz;
  int64_t nl, nn;
  const uint8_t* mask;
  // patch operator maps in the working precision and in FP64
  const PatchMaps* pm;
  const PatchMaps* pm64;
  int Gp, Gp64, Gq;
  // device arrays (workspace slots 48..)
  C *x, *r, *p, *q, *z, *dinv, *w, *winv, *u, *vinv, *part;
  double2 *x64, *b64, *ax, *dp;
  double *kco, *sc, *ring, *sums;
  int *act, *sel;

  void launch_patch(const C* xin, C* yout, bool epi) {
    const R* coef;
    const R* hf;
    if constexpr (sizeof(R) == 4) {
      coef = sys->pt_coef32.p; hf = sys->pt_face32.p;
    } else {
      coef = sys->pt_coef64.p; hf = sys->pt_face64.p;
    }
    const size_t lds = (size_t)(pm->lds + pm->maxfr) * NBL * sizeof(C);
#define SW_PT pm->np, Gp, s->mesh->nc, s->nbf, pm->lds, coef, pm->lp.p, pm->uptr.p, pm->dof.p, \
              pm->seg.p, pm->out.p, pm->fptr.p, s->pt_fmeta.p, pm->fpos.p, hf, \
              s->ref_face_dofs.p, s->ref_face.p, kco, mask, xin, yout, part, dp
    const int cb = pm->cb;
    if (cb == 256) {
      if (epi) SUCEM_LAUNCH((k_sw_patch<R, ORDER, NFD, NB, NBL, 256, true>), Gp, 256, lds, st, SW_PT);
      else SUCEM_LAUNCH((k_sw_patch<R, ORDER, NFD, NB, NBL, 256, false>), Gp, 256, lds, st, SW_PT);
    } else if constexpr (ORDER == 2) {
      if (epi) SUCEM_LAUNCH((k_sw_patch<R, ORDER, NFD, NB, NBL, 192, true>), Gp, 192, lds, st, SW_PT);
      else SUCEM_LAUNCH((k_sw_patch<R, ORDER, NFD, NB, NBL, 192, false>), Gp, 192, lds, st, SW_PT);
    }
#undef SW_PT
    SUCEM_LAUNCH((k_sw_psum<R, NB>), Gq, TB, 0, st, (int)pm->npd, Gq, pm->qptr.p, pm->qdof.p,
                 (const C*)part, mask, yout);
  }

  // FP64 operator on x64 -> ax (all systems): the same kernels in FP64 with the FP64 maps
  void apply_hi() {
    const size_t lds = (size_t)(pm64->lds + pm64->maxfr) * sizeof(double2);
    double2* part64 = reinterpret_cast<double2*>(part);   // sized for the FP64 maps
#define SW_PT64 pm64->np, Gp64, s->mesh->nc, s->nbf, pm64->lds, sys->pt_coef64.p, pm64->lp.p, \
                pm64->uptr.p, pm64->dof.p, pm64->seg.p, pm64->out.p, pm64->fptr.p, \
                s->pt_fmeta.p, pm64->fpos.p, sys->pt_face64.p, s->ref_face_dofs.p, \
                s->ref_face.p, kco, mask, (const double2*)x64, ax, part64, dp
    if (pm64->cb == 256)
      SUCEM_LAUNCH((k_sw_patch<double, ORDER, NFD, NB, 1, 256, false>), Gp64, 256, lds, st, SW_PT64);
    else if constexpr (ORDER == 2)
      SUCEM_LAUNCH((k_sw_patch<double, ORDER, NFD, NB, 1, 192, false>), Gp64, 192, lds, st, SW_PT64);
#undef SW_PT64
    SUCEM_LAUNCH((k_sw_psum<double, NB>), Gq, TB, 0, st, (int)pm64->npd, Gq, pm64->qptr.p,
                 pm64->qdof.p, (const double2*)part64, mask, ax);
  }

  void precondition() {
    const R inv_den = (R)(1.0 / s->g_den);
    const R* xyz;
    if constexpr (sizeof(R) == 4) xyz = s->xyz32.p; else xyz = s->mesh->xyz.p;
    const int nv = (int)s->mesh->nv, nE = (int)s->nE;
    const int Gv = (int)(((int64_t)nv * NB * LV + 8 * TB - 1) / (8 * TB) * 8);
    int Gt = Gv;
    if (ORDER == 2) Gt += (int)(((int64_t)nE * NB * LE + 8 * TB - 1) / (8 * TB) * 8);
    SUCEM_LAUNCH((k_sw_aux_t<R, NB, LV, LE, ORDER>), Gt, TB, 0, st, Gv, nv, nE, inv_den,
                 s->ve_ptr.p, s->ve_ent.p, s->vf_ptr.p, s->vf_ent.p, s->ef_ptr.p, s->ef_ent.p,
                 s->edges.p, s->faces.p, xyz, (const C*)r, (const C*)winv, (const C*)vinv, w, u);
    SUCEM_LAUNCH((k_sw_aux_z<R, NB, ORDER>), Gz, TB, 0, st, n, Gz, nv, nE, inv_den, s->edges.p,
                 s->faces.p, s->face_edges.p, xyz, (const C*)w, (const C*)u, (const C*)dinv,
                 (const C*)r, z, mask, dp);
  }

  // one COCG iteration of every running system; k = slot of |r_j| in the ring
  void iteration(int k) {
    launch_patch(p, q, true);
    SUCEM_LAUNCH((k_sw_alpha<NB>), 1, 64 * NB, 0, st, Gp, (const double2*)dp, (const int*)act, sc);
    SUCEM_LAUNCH((k_sw_xr<R, NB>), G, TB, 0, st, n, G, (const double*)sc, x, r, (const C*)p,
                 (const C*)q);
    precondition();
    SUCEM_LAUNCH((k_sw_beta<NB>), 1, 64 * NB, 0, st, Gz, (const double2*)dp, (const int*)act,
                 (const int*)sel, sc, ring, k, 0);
    SUCEM_LAUNCH((k_sw_pz<R, NB>), G, TB, 0, st, n, G, (const double*)sc, (const C*)z, p);
  }

  // restart the selected systems from their (replaced) residual: z = M^-1 r, rho, p = z
  void restart() {
    precondition();
    SUCEM_LAUNCH((k_sw_beta<NB>), 1, 64 * NB, 0, st, Gz, (const double2*)dp, (const int*)act,
                 (const int*)sel, sc, ring, 0, 1);
    SUCEM_LAUNCH((k_sw_restart_p<R, NB>), G, TB, 0, st, n, G, (const int*)sel, (const C*)z, p);
  }
};

template <typename T>
std::vector<T> d2h_vec(Context* c, hipStream_t st, const T* d, int64_t n) {
  std::vector<T> h(n);
  SUCEM_HIP(hipMemcpyAsync(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost, st));
  SUCEM_HIP(hipStreamSynchronize(st));
  return h;
}

template <typename R, int NB, int ORDER>
void sweep_chunk(Context* c, System* sys, int nb, const double* k0, const double2* B, double2* X,
                 const SolveOpts& o, bool use_graph, SolveStats* stats) {
  using C = typename C2<R>::T;
  using K = Chunk<R, NB, ORDER>;
  const bool lo = sizeof(R) == 4;
  auto t0 = std::chrono::steady_clock::now();
  K k;
  k.c = c;
  k.sys = sys;
  k.s = sys->space;
  const Space* s = sys->space;
  const int n = (int)s->ndofs;
  k.n = n;
  k.mask = o.eliminate_pec ? s->pec.p : nullptr;
  const int cb = patch_cb(ORDER, lo), cb64 = patch_cb(ORDER, false);
  SUCEM_CHECK(pt_prepare(c, sys, cb) && pt_fits<R>(s, cb) && pt_prepare(c, sys, cb64) &&
                  pt_fits<double>(s, cb64),
              SUCEM_ERR_UNSUPPORTED, "the batched sweep needs the patch operator on this space");
  k.pm = &s->pt[cb == 192 ? 1 : 2];
  k.pm64 = &s->pt[cb64 == 192 ? 1 : 2];
  SUCEM_CHECK((int64_t)(k.pm->lds + k.pm->maxfr) * K::NBL * (int64_t)sizeof(C) <=
                  (int64_t)s->pt_max_lds,
              SUCEM_ERR_UNSUPPORTED, "the batched patch kernel needs more LDS than the device has");
  k.Gp = (int)((k.pm->np + 7) / 8 * 8);
  k.Gp64 = (int)((k.pm64->np + 7) / 8 * 8);
  k.Gq = op_grid(c, std::max<int64_t>(std::max(k.pm->npd, k.pm64->npd), 1), 16);
  k.G = op_grid(c, n, 8);
  k.Gz = op_grid(c, n, 4);
  ensure_gradient(c, sys->space);
  ensure_entity_maps(c, sys->space);
  ensure_nodal(c, sys->space);
  k.nl = s->nlag;
  k.nn = 3 * s->mesh->nv;
  const int64_t ne = (int64_t)n * NB;
  const int Gmax = std::max({k.Gp, k.Gp64, k.G, k.Gz});
  const int64_t npart = std::max(k.pm->npe * (int64_t)sizeof(C), k.pm64->npe * 16) * NB;
  // workspace slots 48..79
  k.x = ws_take<C>(c, 48, lo ? ne : 0);
  k.r = ws_take<C>(c, 49, ne);
  k.p = ws_take<C>(c, 50, ne);
  k.q = ws_take<C>(c, 51, ne);
  k.z = ws_take<C>(c, 52, ne);
  k.dinv = ws_take<C>(c, 53, ne);
  k.w = ws_take<C>(c, 54, k.nl * NB);
  k.winv = ws_take<C>(c, 55, k.nl * NB);
  k.u = ws_take<C>(c, 56, k.nn * NB);
  k.vinv = ws_take<C>(c, 57, k.nn * NB);
  k.part = (C*)ws_take<unsigned char>(c, 58, npart);
  k.x64 = ws_take<double2>(c, 59, ne);
  k.b64 = ws_take<double2>(c, 60, ne);
  k.ax = ws_take<double2>(c, 61, ne);
  k.dp = ws_take<double2>(c, 62, (int64_t)2 * NB * Gmax);
  k.kco = ws_take<double>(c, 63, 2 * NB);
  k.sc = ws_take<double>(c, 64, NB * SW);
  k.ring = ws_take<double>(c, 65, (int64_t)o.check_every * NB);
  k.sums = ws_take<double>(c, 66, NB);
  k.act = ws_take<int>(c, 67, NB);
  k.sel = ws_take<int>(c, 68, NB);
  if (!lo) k.x = reinterpret_cast<C*>(k.x64);
  // per-system weights (the single-system set-up routines, one system at a time, into
  // a scratch column, then interleaved)
  std::vector<double> kco(2 * NB);
  for (int j = 0; j < NB; ++j) {
    const double kk = k0[j < nb ? j : 0];
    kco[j] = -kk * kk;
    kco[NB + j] = (sys->which & SUCEM_MAT_S0) ? kk : 0.0;
  }
  h2d(c, k.kco, kco.data(), kco.size());
  {
    C* tmp = ws_take<C>(c, 69, std::max<int64_t>({(int64_t)n, k.nl, k.nn}));
    const bool masked = k.mask != nullptr;
    for (int j = 0; j < NB; ++j) {
      const Coef cf = driven_coef(k0[j < nb ? j : 0]);
      Coef cj = cf;
      if (!(sys->which & SUCEM_MAT_S0)) cj.s0 = 0;
      cj.ti = -std::max(0.0, o.precond_shift) * cj.t;   // the preconditioner's complex shift
      sys_dinv<R>(c, sys, cj, k.mask, tmp);
      SUCEM_LAUNCH((k_sw_put<C>), grid_for(n, TB), TB, 0, c->stream, (int64_t)n, NB, j,
                   (const C*)tmp, k.dinv);
      grad_winv<R>(c, sys, cj, masked, true, tmp);
      SUCEM_LAUNCH((k_sw_put<C>), grid_for(k.nl, TB), TB, 0, c->stream, k.nl, NB, j,
                   (const C*)tmp, k.winv);
      nodal_winv<R>(c, sys, cj, masked, tmp);
      SUCEM_LAUNCH((k_sw_put<C>), grid_for(k.nn, TB), TB, 0, c->stream, k.nn, NB, j,
                   (const C*)tmp, k.vinv);
    }
  }
  SUCEM_LAUNCH((k_sw_interleave<double2>), grid_for(ne, TB), TB, 0, c->stream, (int64_t)n, NB, nb,
               B, k.b64);
  // the iteration runs on the library stream, ordered after the set-up
  k.st = c->graph_stream();
  hipEvent_t ev;
  SUCEM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  SUCEM_HIP(hipEventRecord(ev, c->stream));
  SUCEM_HIP(hipStreamWaitEvent(k.st, ev, 0));
  hipStream_t st = k.st;

  SUCEM_LAUNCH((k_sw_norm<NB>), k.G, TB, 0, st, n, k.G, (const double2*)k.b64, k.mask, k.dp);
  SUCEM_LAUNCH((k_sw_sum<NB>), 1, 64 * NB, 0, st, k.G, (const double2*)k.dp, k.sums);
  std::vector<double> bn = d2h_vec(c, st, k.sums, NB);
  for (auto& v : bn) v = std::sqrt(v);
  std::vector<int> act(NB, 0), sel(NB, 0);
  for (int j = 0; j < nb; ++j) act[j] = bn[j] > 0.0 ? 1 : 0;
  auto upload = [&](int* d, const std::vector<int>& h) {
    SUCEM_HIP(hipMemcpyAsync(d, h.data(), sizeof(int) * NB, hipMemcpyHostToDevice, st));
    SUCEM_HIP(hipStreamSynchronize(st));
  };
  // x = 0, r = b, start every running system
  auto zero = [&](void* a, int64_t bytes) {
    const int64_t words = bytes / 4;
    SUCEM_LAUNCH(k_sw_zero, grid_for(words, TB), TB, 0, st, words, (uint32_t*)a);
  };
  zero(k.x64, ne * (int64_t)sizeof(double2));
  if (lo) zero(k.x, ne * (int64_t)sizeof(C));
  zero(k.sc, (int64_t)NB * SW * (int64_t)sizeof(double));
  SUCEM_LAUNCH((k_sw_to_lo<R>), grid_for(ne, TB), TB, 0, st, ne, (const double2*)k.b64, k.r);
  upload(k.act, act);
  upload(k.sel, act);
  k.restart();

  std::vector<double> rel(NB, 1.0), last_true(NB, 1.0), best(NB, 1.0);
  std::vector<int> done(NB, 0), its(NB, 0), ru(NB, 0), flat(NB, 0), stag(NB, 0);
  // FP32 recursions that diverged or stopped lowering the true residual leave the batch
  // and are finished in FP64 by the single-system solver from their accepted x
  std::vector<int> fail(NB, 0), drop(NB, 0);
  std::vector<int> cyc_start(NB, 0), longest(NB, 0);
  for (int j = 0; j < NB; ++j) if (!act[j]) done[j] = 1;
  const int ce = o.check_every;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int it = 0;
  auto all_done = [&] { return std::all_of(done.begin(), done.end(), [](int d) { return d != 0; }); };
  try {
    while (it < o.max_iter && !all_done()) {
      const int steps = std::min(ce, o.max_iter - it);
      if (use_graph && steps == ce) {
        if (!exec) {
          SUCEM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
          for (int i = 0; i < ce; ++i) k.iteration(i);
          SUCEM_HIP(hipStreamEndCapture(st, &graph));
          SUCEM_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        }
        SUCEM_HIP(hipGraphLaunch(exec, st));
      } else {
        for (int i = 0; i < steps; ++i) k.iteration(i);
      }
      it += steps;
      std::vector<double> ring = d2h_vec(c, st, k.ring, (int64_t)steps * NB);
      bool any = false, any_drop = false;
      for (int j = 0; j < NB; ++j) {
        sel[j] = 0;
        drop[j] = 0;
        if (done[j]) continue;
        rel[j] = ring[(int64_t)(steps - 1) * NB + j] / bn[j];
        if (lo && !(rel[j] <= 1e3 * last_true[j])) {   // diverged (or not finite) in FP32
          done[j] = fail[j] = drop[j] = 1;
          its[j] = it;
          any_drop = true;
          continue;
        }
        if (!std::isfinite(rel[j])) {   // breakdown of this recursion
          done[j] = 1;
          its[j] = it;
          stag[j] = -1;
          continue;
        }
        const int cyc = it - cyc_start[j];
        const int cap = longest[j] > 0 ? std::max(64, 3 * longest[j])
                                       : std::max(1000, o.max_iter / 8);
        const bool cut = lo && cyc >= cap;
        if (rel[j] <= o.tol || (lo && rel[j] <= o.ru_delta * last_true[j]) || cut) {
          sel[j] = 1;
          any = true;
          if (!cut) longest[j] = std::max(longest[j], cyc);
          cyc_start[j] = it;
        }
      }
      for (int j = 0; j < NB; ++j) act[j] = done[j] ? 0 : 1;
      if (any_drop) {   // their current cycle's correction is discarded
        upload(k.sel, drop);
        SUCEM_LAUNCH((k_sw_accum<R, NB>), k.G, TB, 0, st, n, k.G, (const int*)k.sel, k.x, k.x64,
                     0.0, 1);
      }
      if (!any) {
        upload(k.act, act);
        continue;
      }
      // residual replacement of the selected systems with the FP64 operator
      upload(k.sel, sel);
      if (lo)
        SUCEM_LAUNCH((k_sw_accum<R, NB>), k.G, TB, 0, st, n, k.G, (const int*)k.sel, k.x, k.x64,
                     1.0, 0);
      k.apply_hi();
      SUCEM_LAUNCH((k_sw_resid<R, NB>), k.G, TB, 0, st, n, k.G, (const int*)k.sel,
                   (const double2*)k.b64, (const double2*)k.ax, k.mask, k.r, k.dp);
      SUCEM_LAUNCH((k_sw_sum<NB>), 1, 64 * NB, 0, st, k.G, (const double2*)k.dp, k.sums);
      std::vector<double> tr = d2h_vec(c, st, k.sums, NB);
      bool restart_any = false, any_revert = false;
      std::vector<int> revert(NB, 0), keep(NB, 0);
      for (int j = 0; j < NB; ++j) {
        if (!sel[j]) continue;
        ru[j] += 1;
        const double t = std::sqrt(tr[j]) / bn[j];
        if (lo && !(t < last_true[j])) {
          // the FP32 cycle did not lower the true residual: its correction comes out again
          revert[j] = 1;
          any_revert = true;
          done[j] = fail[j] = 1;
          its[j] = it;
          sel[j] = 0;
          continue;
        }
        keep[j] = 1;
        rel[j] = t;
        last_true[j] = t;
        if (t <= o.tol) {
          done[j] = 1;
          its[j] = it;
          sel[j] = 0;
          continue;
        }
        if (t <= 0.5 * best[j]) {
          best[j] = t;
          flat[j] = 0;
        } else if (++flat[j] >= 3) {   // at the floor of its true residual
          done[j] = 1;
          its[j] = it;
          sel[j] = 0;
          if (lo) fail[j] = 1;   // FP64 may still get further
          else stag[j] = 1;
          continue;
        }
        restart_any = true;
      }
      if (lo) {
        if (any_revert) {
          upload(k.sel, revert);
          SUCEM_LAUNCH((k_sw_accum<R, NB>), k.G, TB, 0, st, n, k.G, (const int*)k.sel, k.x,
                       k.x64, -1.0, 1);
        }
        upload(k.sel, keep);
        SUCEM_LAUNCH((k_sw_accum<R, NB>), k.G, TB, 0, st, n, k.G, (const int*)k.sel, k.x, k.x64,
                     0.0, 1);
      }
      for (int j = 0; j < NB; ++j) act[j] = done[j] ? 0 : 1;
      upload(k.act, act);
      if (restart_any) {
        upload(k.sel, sel);
        k.restart();
      }
    }
  } catch (...) {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipEventDestroy(ev);
    throw;
  }
  if (exec) SUCEM_HIP(hipGraphExecDestroy(exec));
  if (graph) SUCEM_HIP(hipGraphDestroy(graph));
  // systems stopped by the iteration limit: fold in their last correction and report the
  // true residual
  std::vector<int> open(NB, 0);
  bool any_open = false;
  for (int j = 0; j < nb; ++j)
    if (!done[j]) { open[j] = 1; any_open = true; its[j] = it; }
  if (any_open) {
    upload(k.sel, open);
    if (lo)
      SUCEM_LAUNCH((k_sw_accum<R, NB>), k.G, TB, 0, st, n, k.G, (const int*)k.sel, k.x, k.x64,
                   1.0, 1);
    k.apply_hi();
    SUCEM_LAUNCH((k_sw_resid<R, NB>), k.G, TB, 0, st, n, k.G, (const int*)k.sel,
                 (const double2*)k.b64, (const double2*)k.ax, k.mask, k.r, k.dp);
    SUCEM_LAUNCH((k_sw_sum<NB>), 1, 64 * NB, 0, st, k.G, (const double2*)k.dp, k.sums);
    std::vector<double> tr = d2h_vec(c, st, k.sums, NB);
    for (int j = 0; j < nb; ++j)
      if (open[j]) rel[j] = std::sqrt(tr[j]) / bn[j];
  }
  SUCEM_LAUNCH(k_sw_deinterleave, grid_for(ne, TB), TB, 0, st, (int64_t)n, NB, nb,
               (const double2*)k.x64, X);
  SUCEM_HIP(hipEventRecord(ev, st));
  SUCEM_HIP(hipStreamWaitEvent(c->stream, ev, 0));
  SUCEM_HIP(hipStreamSynchronize(st));
  SUCEM_HIP(hipEventDestroy(ev));
  // systems whose FP32 recursion failed: the single-system FP64 solver from their x
  std::vector<SolveStats> fin(NB);
  for (int j = 0; j < nb; ++j) {
    if (!fail[j]) continue;
    SolveOpts o64 = o;
    o64.precision = 64;
    o64.warm = 1;
    o64.max_iter = std::max(1, o.max_iter - its[j]);
    Coef cj = driven_coef(k0[j]);
    if (!(sys->which & SUCEM_MAT_S0)) cj.s0 = 0;
    solve_system(c, sys, cj, B + (int64_t)j * n, X + (int64_t)j * n, o64, fin[j], nullptr, 0);
  }
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (int j = 0; j < nb; ++j) {
    SolveStats& q = stats[j];
    q = SolveStats();
    const bool ok = done[j] && stag[j] == 0 && bn[j] > 0.0 ? rel[j] <= o.tol : false;
    if (bn[j] == 0.0) {
      q.status = 0;
      q.rel_residual = 0;
    } else {
      q.status = stag[j] < 0 ? -1 : (ok ? 0 : std::max(its[j], 1));
      q.rel_residual = rel[j];
    }
    q.stagnated = stag[j] > 0 ? 1 : 0;
    q.iterations = its[j];
    q.reliable_updates = ru[j];
    q.elapsed_ms = ms;
    q.spmv_lo = it;
    q.spmv_hi = ru[j];
    if (fail[j]) {
      q.status = fin[j].status > 0 ? its[j] + fin[j].iterations : fin[j].status;
      q.rel_residual = fin[j].rel_residual;
      q.stagnated = fin[j].stagnated;
      q.iterations = its[j] + fin[j].iterations;
      q.reliable_updates += fin[j].reliable_updates;
      q.spmv_hi += fin[j].spmv_hi;
      q.fp32_failed = 1;
    }
  }
}

template <typename R, int ORDER>
void sweep_dispatch(Context* c, System* sys, int64_t nb, const double* k0, const double2* B,
                    double2* X, const SolveOpts& o, int width, bool use_graph, SolveStats* st) {
  const int64_t n = sys->space->ndofs;
  // systems off .. off + m - 1 as one group of the next width up from m
  auto group = [&](int64_t off, int m) {
    const double2* Bc = B + off * n;
    double2* Xc = X + off * n;
    const double* kc = k0 + off;
    SolveStats* sc = st + off;
    if (m > 4) sweep_chunk<R, 8, ORDER>(c, sys, m, kc, Bc, Xc, o, use_graph, sc);
    else if (m > 2) sweep_chunk<R, 4, ORDER>(c, sys, m, kc, Bc, Xc, o, use_graph, sc);
    else if (m > 1) sweep_chunk<R, 2, ORDER>(c, sys, m, kc, Bc, Xc, o, use_graph, sc);
    else sweep_chunk<R, 1, ORDER>(c, sys, m, kc, Bc, Xc, o, use_graph, sc);
    for (int j = 0; j < m; ++j) sc[j].batch_width = m;
  };
  int64_t off = 0;
  if (width == 0) {
    // Autotuning: the first systems are solved in groups of 2, then 4, then 8 (real solves,
    // each group timed on the host around its synchronised chunk), and the width with the
    // least time per system-iteration -- the frozen lanes of systems that converged early
    // included -- takes the rest.  The first group also pays the first-use set-up, which
    // biases against the narrowest width only.
    width = 8;
    double best = 1e300;
    for (int w : {2, 4, 8}) {
      if (off + w > nb) break;
      const auto t0 = std::chrono::steady_clock::now();
      group(off, w);
      const double sec =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      int64_t its = 0;
      for (int j = 0; j < w; ++j) its += st[off + j].iterations;
      const double per = sec / (double)std::max<int64_t>(its, 1);
      if (per < best) {
        best = per;
        width = w;
      }
      off += w;
    }
  }
  for (; off < nb; off += width) group(off, (int)std::min<int64_t>(width, nb - off));
}

}  // namespace

bool sys_supports_sweep(Context* c, System* sys) {
  const Space* s = sys->space;
  if (s->order != 1 && s->order != 2) return false;
  const int cb32 = patch_cb(s->order, true), cb64 = patch_cb(s->order, false);
  return pt_prepare(c, sys, cb32) && pt_fits<float>(s, cb32) && pt_prepare(c, sys, cb64) &&
         pt_fits<double>(s, cb64);
}

void solve_system_sweep(Context* c, System* sys, int64_t nb, const double* k0, const double2* B,
                        double2* X, const SolveOpts& o, int width, bool use_graph,
                        SolveStats* st) {
  SUCEM_CHECK(sys->st64.p, SUCEM_ERR_INVALID, "S and T must be assembled");
  SUCEM_CHECK(o.method == SUCEM_METHOD_COCG && o.precond == SUCEM_PRECOND_AUXILIARY,
              SUCEM_ERR_UNSUPPORTED,
              "the batched sweep runs COCG with the auxiliary-space preconditioner");
  SUCEM_CHECK(o.precision == 32 || o.precision == 64, SUCEM_ERR_INVALID,
              "precision must be 32 or 64");
  SUCEM_CHECK(o.max_iter > 0 && o.check_every > 0 && o.check_every <= 256, SUCEM_ERR_INVALID,
              "bad iteration limits (check_every at most 256)");
  SUCEM_CHECK(width == 0 || width == 1 || width == 2 || width == 4 || width == 8,
              SUCEM_ERR_INVALID, "batch width must be 1, 2, 4 or 8 (0: autotuned)");
  const int order = sys->space->order;
  SUCEM_CHECK(order == 1 || order == 2, SUCEM_ERR_UNSUPPORTED, "orders 1 and 2 only");
  if (o.precision == 32) {
    if (order == 1) sweep_dispatch<float, 1>(c, sys, nb, k0, B, X, o, width, use_graph, st);
    else sweep_dispatch<float, 2>(c, sys, nb, k0, B, X, o, width, use_graph, st);
  } else {
    if (order == 1) sweep_dispatch<double, 1>(c, sys, nb, k0, B, X, o, width, use_graph, st);
    else sweep_dispatch<double, 2>(c, sys, nb, k0, B, X, o, width, use_graph, st);
  }
}

}  // namespace sucem
