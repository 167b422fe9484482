// Batched driven solves: NB systems A_j x_j = b_j on one mesh with
// A_j = S - k0_j^2 T + i k0_j S0 (frequency sweeps, many right-hand sides).
//
// Vectors are interleaved [row][system], so one column gather of the operator
// reads a contiguous segment holding that entry of every system, and the
// matrix (column indices and (S, T) pairs) is read once per iteration for the
// whole batch.  Each system runs its own preconditioned COCG (Jacobi + gradient
// correction) with its own scalars; per-system reductions are finished by a
// one-block kernel, as in the single-system solver.  FP64 throughout.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "internal.hpp"
#include "kernels.hpp"
#include "ops.hpp"

namespace sucem {
namespace {

constexpr int TB = 256;
using C = double2;

// ---- per-system reduction ------------------------------------------------------
__device__ __forceinline__ void finalize_sys(const RedSite& rs, int NB, int j, double2 s0,
                                             double2 s1) {
  RedSite one = rs;
  one.scal = rs.scal + j * S_COUNT;
  one.hist = nullptr;
  finalize(one, s0, s1);
  if (rs.hist && (rs.mode == FIN_COCG_BETA)) rs.hist[(int64_t)rs.it * NB + j] = sqrt(s1.x);
}

// Every thread holds values for system j = threadIdx.x % NB.  Block sums per system
// -> partials part[lb*NB + j] (first G*NB) and part[(G + lb)*NB + j]; k_finalize_nb,
// launched after the producing kernel, sums them per system in block order and
// finalises each system's scalars (stream order, no device-scope fences).
template <int NB>
__device__ void reduce_finalize_nb(const RedSite& rs, int G, double2 v0, double2 v1) {
  __shared__ double2 s0[TB], s1[TB];
  const int t = threadIdx.x;
  s0[t] = v0;
  s1[t] = v1;
  __syncthreads();
  for (int step = TB / 2; step >= NB; step >>= 1) {
    if (t < step) {
      s0[t].x += s0[t + step].x; s0[t].y += s0[t + step].y;
      s1[t].x += s1[t + step].x; s1[t].y += s1[t + step].y;
    }
    __syncthreads();
  }
  const int lb = xcd_block(G);
  if (t < NB) {
    rs.part[(int64_t)lb * NB + t] = s0[t];
    rs.part[(int64_t)(G + lb) * NB + t] = s1[t];
  }
}

template <int NB>
__global__ __launch_bounds__(TB) void k_finalize_nb(RedSite rs, int G) {
  __shared__ double2 s0[TB], s1[TB];
  const int t = threadIdx.x;
  const int j = t % NB;
  double2 a = make_double2(0, 0), b = make_double2(0, 0);
  for (int g = t / NB; g < G; g += TB / NB) {
    const double2 p = rs.part[(int64_t)g * NB + j], q = rs.part[(int64_t)(G + g) * NB + j];
    a.x += p.x; a.y += p.y;
    b.x += q.x; b.y += q.y;
  }
  s0[t] = a;
  s1[t] = b;
  __syncthreads();
  for (int step = TB / 2; step >= NB; step >>= 1) {
    if (t < step) {
      s0[t].x += s0[t + step].x; s0[t].y += s0[t + step].y;
      s1[t].x += s1[t + step].x; s1[t].y += s1[t + step].y;
    }
    __syncthreads();
  }
  if (t < NB) finalize_sys(rs, NB, t, s0[t], s1[t]);
}

struct Rows { int start, end; };
__device__ __forceinline__ Rows rows_of(int n, int G) {
  const int lb = xcd_block(G);
  const int chunk = (n + G - 1) / G;
  Rows r;
  r.start = lb * chunk;
  r.end = min(n, r.start + chunk);
  return r;
}

// ---- batched operator -----------------------------------------------------------
// y_j = cs S x_j + ct_j T x_j + i c0_j S0 x_j ; coef = {cs, ct[NB], c0[NB]}.
// Lanes of a row group: (system jv = lane % NB, nonzero stride lr = lane / NB).
template <int NB, int LR, int EPI>
__global__ __launch_bounds__(TB) void k_spmm(int n, int G, const int32_t* __restrict__ rp,
                                             const int32_t* __restrict__ col,
                                             const double2* __restrict__ st,
                                             const int32_t* __restrict__ s0p,
                                             const int32_t* __restrict__ s0c,
                                             const double* __restrict__ s0v,
                                             const double* __restrict__ coef,
                                             const uint8_t* __restrict__ mask,
                                             const C* __restrict__ x, C* __restrict__ y,
                                             const C* __restrict__ w, RedSite rs) {
  constexpr int L = NB * LR;
  constexpr int RPB = TB / L;
  const int lane = threadIdx.x % L;
  const int jv = lane % NB, lr = lane / NB;
  const double cs = coef[0], ct = coef[1 + jv], c0 = coef[1 + NB + jv];
  const int lb = xcd_block(G);
  int chunk = (n + G - 1) / G;
  chunk = (chunk + RPB - 1) / RPB * RPB;
  const int start = lb * chunk, end = min(n, start + chunk);
  double2 d1 = make_double2(0, 0), d2 = make_double2(0, 0);
  for (int base = start; base < end; base += RPB) {
    const int r = base + threadIdx.x / L;
    double2 as = make_double2(0, 0), at = make_double2(0, 0), a0 = make_double2(0, 0);
    if (r < end) {
      const int k1 = rp[r + 1];
      for (int k = rp[r] + lr; k < k1; k += LR) {
        const double2 v = st[k];
        const C xc = x[(int64_t)col[k] * NB + jv];
        as.x += v.x * xc.x; as.y += v.x * xc.y;
        at.x += v.y * xc.x; at.y += v.y * xc.y;
      }
      if (s0p) {
        const int q1 = s0p[r + 1];
        for (int q = s0p[r] + lr; q < q1; q += LR) {
          const double v = s0v[q];
          const C xc = x[(int64_t)s0c[q] * NB + jv];
          a0.x += v * xc.x; a0.y += v * xc.y;
        }
      }
    }
#pragma unroll
    for (int o = L / 2; o >= NB; o >>= 1) {
      as.x += __shfl_xor(as.x, o); as.y += __shfl_xor(as.y, o);
      at.x += __shfl_xor(at.x, o); at.y += __shfl_xor(at.y, o);
      a0.x += __shfl_xor(a0.x, o); a0.y += __shfl_xor(a0.y, o);
    }
    if (lr == 0 && r < end) {
      C yr = make_double2(cs * as.x + ct * at.x - c0 * a0.y, cs * as.y + ct * at.y + c0 * a0.x);
      if (mask && mask[r]) yr = make_double2(0, 0);
      y[(int64_t)r * NB + jv] = yr;
      if (EPI) {
        const C wr = w[(int64_t)r * NB + jv];
        d1.x += wr.x * yr.x - wr.y * yr.y;
        d1.y += wr.x * yr.y + wr.y * yr.x;
        d2.x += yr.x * yr.x + yr.y * yr.y;
      }
    }
  }
  if (EPI) reduce_finalize_nb<NB>(rs, G, d1, d2);
}

// ---- batched vector kernels (element e = row*NB + system) ------------------------
template <int NB>
__global__ __launch_bounds__(TB) void k_bnorm(int n, int G, const C* b, const uint8_t* mask,
                                              RedSite rs) {
  const Rows rr = rows_of(n, G);
  double2 s = make_double2(0, 0);
  for (int64_t e = (int64_t)rr.start * NB + threadIdx.x; e < (int64_t)rr.end * NB; e += TB)
    if (!(mask && mask[e / NB])) s.x += b[e].x * b[e].x + b[e].y * b[e].y;
  reduce_finalize_nb<NB>(rs, G, s, make_double2(0, 0));
}

// r = b - ax (masked rows zero); reduces |r|^2 per system
template <int NB>
__global__ __launch_bounds__(TB) void k_bresid(int n, int G, const C* b, const C* ax,
                                               const uint8_t* mask, C* r, RedSite rs) {
  const Rows rr = rows_of(n, G);
  double2 s = make_double2(0, 0);
  for (int64_t e = (int64_t)rr.start * NB + threadIdx.x; e < (int64_t)rr.end * NB; e += TB) {
    C v = make_double2(b[e].x - ax[e].x, b[e].y - ax[e].y);
    if (mask && mask[e / NB]) v = make_double2(0, 0);
    r[e] = v;
    s.x += v.x * v.x + v.y * v.y;
  }
  reduce_finalize_nb<NB>(rs, G, s, make_double2(0, 0));
}

// Systems with act[j] == 0 (converged or broken down) are left as they are, so the
// returned x_j is the one whose residual was checked.  Element e = row*NB + j with TB a
// multiple of NB, so a thread always handles the same system.
template <int NB>
__global__ void k_bxr(int n, int G, const double* scal, const int* act, C* x, C* r, const C* p,
                      const C* q) {
  const Rows rr = rows_of(n, G);
  const int j = threadIdx.x % NB;
  if (!act[j]) return;
  const C a = scal2(scal + j * S_COUNT, S_ALPHA);
  for (int64_t e = (int64_t)rr.start * NB + threadIdx.x; e < (int64_t)rr.end * NB; e += TB) {
    const C ap = cmul(a, p[e]), aq = cmul(a, q[e]);
    x[e].x += ap.x; x[e].y += ap.y;
    r[e].x -= aq.x; r[e].y -= aq.y;
  }
}

template <int NB>
__global__ void k_bpz(int n, int G, const double* scal, const int* act, const C* z, C* p) {
  const Rows rr = rows_of(n, G);
  const int j = threadIdx.x % NB;
  if (!act[j]) return;
  const C b = scal2(scal + j * S_COUNT, S_BETA);
  for (int64_t e = (int64_t)rr.start * NB + threadIdx.x; e < (int64_t)rr.end * NB; e += TB) {
    const C bp = cmul(b, p[e]);
    p[e] = make_double2(z[e].x + bp.x, z[e].y + bp.y);
  }
}

template <int NB>
__global__ void k_bcopy(int n, int G, const C* a, C* b) {
  const Rows rr = rows_of(n, G);
  for (int64_t e = (int64_t)rr.start * NB + threadIdx.x; e < (int64_t)rr.end * NB; e += TB)
    b[e] = a[e];
}

// dinv[row*NB + j] = 1 / (cs S_ii + ct_j T_ii + i c0_j S0_ii)
template <int NB>
__global__ void k_bdinv(int64_t n, const int32_t* diag, const double2* st, const int32_t* s0diag,
                        const double* s0v, const double* coef, const uint8_t* mask, C* dinv) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * NB;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / NB;
    const int j = (int)(e % NB);
    const double2 v = st[diag[r]];
    const double re = coef[0] * v.x + coef[1 + j] * v.y;
    const double im = (s0diag && s0diag[r] >= 0) ? coef[1 + NB + j] * s0v[s0diag[r]] : 0.0;
    const double den = re * re + im * im;
    dinv[e] = ((mask && mask[r]) || den == 0.0) ? make_double2(0, 0)
                                                 : make_double2(re / den, -im / den);
  }
}

template <int NB>
__global__ void k_bwinv(int64_t nlag, const double* d3, const double* coef, double theta,
                        C* winv) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nlag * NB;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = e / NB;
    const int j = (int)(e % NB);
    const double re = coef[0] * d3[3 * v] + coef[1 + j] * d3[3 * v + 1];
    const double im = coef[1 + NB + j] * d3[3 * v + 2];
    const double den = re * re + im * im;
    winv[e] = den > 0.0 ? make_double2(theta * re / den, -theta * im / den) : make_double2(0, 0);
  }
}

// w[node*NB + j] = winv .* (G^T r_j)
template <int NB, int LR>
__global__ __launch_bounds__(TB) void k_bgrad_t(int64_t nlag, const int32_t* gt_ptr,
                                                const int32_t* gt_col, const double* gt_val,
                                                const C* r, const C* winv, C* w) {
  constexpr int L = NB * LR;
  const int lane = threadIdx.x % L;
  const int jv = lane % NB, lr = lane / NB;
  const int64_t group = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / L;
  const int64_t ngroups = ((int64_t)gridDim.x * blockDim.x) / L;
  for (int64_t v = group; v < nlag; v += ngroups) {
    double2 acc = make_double2(0, 0);
    for (int q = gt_ptr[v] + lr; q < gt_ptr[v + 1]; q += LR) {
      const double g = gt_val[q];
      const C x = r[(int64_t)gt_col[q] * NB + jv];
      acc.x += g * x.x;
      acc.y += g * x.y;
    }
#pragma unroll
    for (int o = L / 2; o >= NB; o >>= 1) {
      acc.x += __shfl_xor(acc.x, o);
      acc.y += __shfl_xor(acc.y, o);
    }
    if (lr == 0) w[v * NB + jv] = cmul(winv[v * NB + jv], acc);
  }
}

// z = dinv .* r + G w ; reduces r.z and |r|^2 per system
template <int NB>
__global__ __launch_bounds__(TB) void k_bgrad_z(int n, int G, const int32_t* g_ptr,
                                                const int32_t* g_col, const double* g_val,
                                                const C* w, const C* dinv, const C* r, C* z,
                                                const uint8_t* mask, RedSite rs) {
  const Rows rr = rows_of(n, G);
  double2 d1 = make_double2(0, 0), d2 = make_double2(0, 0);
  for (int64_t e = (int64_t)rr.start * NB + threadIdx.x; e < (int64_t)rr.end * NB; e += TB) {
    const int64_t i = e / NB;
    const int j = (int)(e % NB);
    const C ri = r[e];
    C zi = cmul(dinv[e], ri);
    for (int q = g_ptr[i]; q < g_ptr[i + 1]; ++q) {
      const double g = g_val[q];
      const C v = w[(int64_t)g_col[q] * NB + j];
      zi.x += g * v.x;
      zi.y += g * v.y;
    }
    if (mask && mask[i]) zi = make_double2(0, 0);
    z[e] = zi;
    d1.x += ri.x * zi.x - ri.y * zi.y;
    d1.y += ri.x * zi.y + ri.y * zi.x;
    d2.x += ri.x * ri.x + ri.y * ri.y;
  }
  reduce_finalize_nb<NB>(rs, G, d1, d2);
}

// interleave: out[row*NB + j] = in[j*n + row] for j < nb (zero for padding systems)
template <int NB>
__global__ void k_interleave(int64_t n, int nb, const C* in, C* out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * NB;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / NB;
    const int j = (int)(e % NB);
    out[e] = j < nb ? in[(int64_t)j * n + r] : make_double2(0, 0);
  }
}

template <int NB>
__global__ void k_deinterleave(int64_t n, int nb, const C* in, C* out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * NB;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / NB;
    const int j = (int)(e % NB);
    if (j < nb) out[(int64_t)j * n + r] = in[e];
  }
}

__global__ void k_bzero_u32(int n, unsigned* p) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

__global__ void k_bzero_f64(int64_t n, double* p) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

template <int NB>
void solve_chunk(Context* c, System* sys, int nb, const double* k0, const double2* B, double2* X,
                 const SolveOpts& o, SolveStats* st) {
  constexpr int LR = NB >= 16 ? 4 : (NB >= 8 ? 4 : 8);
  const Space* s = sys->space;
  const int n = (int)s->ndofs;
  const uint8_t* mask = o.eliminate_pec ? s->pec.p : nullptr;
  ensure_gradient(c, sys->space);
  const int64_t nl = s->nlag;
  const double* d3 = gradient_diag(c, sys, mask != nullptr);
  const int G = op_grid(c, n, 8);
  const bool has_s0 = (sys->which & SUCEM_MAT_S0) && s->nnz0 > 0;
  auto t0 = std::chrono::steady_clock::now();

  // coefficients {cs, ct[NB], c0[NB]} (padding systems get k0 = 1 and a zero right-hand side)
  std::vector<double> coef(1 + 2 * NB, 0.0);
  coef[0] = 1.0;
  for (int j = 0; j < NB; ++j) {
    const double kk = j < nb ? k0[j] : 1.0;
    coef[1 + j] = -kk * kk;
    coef[1 + NB + j] = kk;
  }
  DBuf<double> coef_d(coef.size());
  h2d(c, coef_d.p, coef.data(), coef.size());
  const int64_t ne = (int64_t)n * NB;
  DBuf<C> b(ne), x(ne), r(ne), p(ne), q(ne), z(ne), dinv(ne), ax(ne);
  DBuf<C> w(nl * NB), winv(nl * NB);
  const int sites = 4;
  DBuf<double2> part((int64_t)sites * 2 * G * NB);
  DBuf<unsigned> counters(sites);
  DBuf<double> scal((int64_t)NB * S_COUNT);
  DBuf<double> hist((int64_t)(o.max_iter + 1) * NB);
  SUCEM_LAUNCH(k_bzero_u32, 1, 64, 0, c->stream, sites, counters.p);
  SUCEM_LAUNCH(k_bzero_f64, grid_for((int64_t)NB * S_COUNT, TB), TB, 0, c->stream,
               (int64_t)NB * S_COUNT, scal.p);
  SUCEM_LAUNCH((k_bzero_f64), grid_for(ne * 2, TB), TB, 0, c->stream, ne * 2, (double*)x.p);
  auto site = [&](int k, int mode, int it = 0) {
    RedSite rs;
    rs.part = part.p + (int64_t)k * 2 * G * NB;
    rs.counter = counters.p + k;
    rs.scal = scal.p;
    rs.hist = hist.p;
    rs.mode = mode;
    rs.it = it;
    return rs;
  };
  SUCEM_LAUNCH((k_interleave<NB>), grid_for(ne, TB), TB, 0, c->stream, (int64_t)n, nb, B, b.p);
  SUCEM_LAUNCH((k_bdinv<NB>), grid_for(ne, TB), TB, 0, c->stream, (int64_t)n, s->diag.p, sys->st64.p,
               has_s0 ? s->s0_diag.p : nullptr, sys->s0_64.p, coef_d.p, mask, dinv.p);
  SUCEM_LAUNCH((k_bwinv<NB>), grid_for(nl * NB, TB), TB, 0, c->stream, nl, d3, coef_d.p,
               gradient_weight(s->order, false), winv.p);
  auto finish = [&](const RedSite& rs) {
    SUCEM_LAUNCH((k_finalize_nb<NB>), 1, TB, 0, c->stream, rs, G);
  };
  auto precondition = [&](const RedSite& rs) {
    SUCEM_LAUNCH((k_bgrad_t<NB, LR>), grid_for(nl * NB * LR, TB, (int64_t)c->num_cus * 16), TB, 0,
                 c->stream, nl, s->gt_ptr.p, s->gt_col.p, s->gt_val.p, r.p, winv.p, w.p);
    SUCEM_LAUNCH((k_bgrad_z<NB>), G, TB, 0, c->stream, n, G, s->g_ptr.p, s->g_col.p, s->g_val.p,
                 w.p, dinv.p, r.p, z.p, mask, rs);
    finish(rs);
  };
  auto spmm = [&](const C* xin, C* yout, const C* wv, int epi, const RedSite& rs) {
    if (epi) {
      SUCEM_LAUNCH((k_spmm<NB, LR, 1>), G, TB, 0, c->stream, n, G, s->row_ptr.p, s->col.p,
                   sys->st64.p, has_s0 ? s->s0_ptr.p : nullptr, s->s0_col.p, sys->s0_64.p,
                   coef_d.p, mask, xin, yout, wv, rs);
      finish(rs);
    } else
      SUCEM_LAUNCH((k_spmm<NB, LR, 0>), G, TB, 0, c->stream, n, G, s->row_ptr.p, s->col.p,
                   sys->st64.p, has_s0 ? s->s0_ptr.p : nullptr, s->s0_col.p, sys->s0_64.p,
                   coef_d.p, mask, xin, yout, wv, rs);
  };
  auto read = [&](const double* d, int64_t cnt) {
    std::vector<double> h(cnt);
    SUCEM_HIP(hipMemcpyAsync(h.data(), d, sizeof(double) * cnt, hipMemcpyDeviceToHost, c->stream));
    c->sync();
    return h;
  };
  auto scal_slot = [&](int slot) {
    std::vector<double> all = read(scal.p, (int64_t)NB * S_COUNT);
    std::vector<double> v(NB);
    for (int j = 0; j < NB; ++j) v[j] = all[j * S_COUNT + slot];
    return v;
  };

  SUCEM_LAUNCH((k_bnorm<NB>), G, TB, 0, c->stream, n, G, b.p, mask, site(0, FIN_NORM));
  finish(site(0, FIN_NORM));
  std::vector<double> bn = scal_slot(S_BNORM);
  for (auto& v : bn) v = std::sqrt(v);
  // r = b (x0 = 0) ; z = M r ; p = z
  SUCEM_LAUNCH((k_bcopy<NB>), G, TB, 0, c->stream, n, G, b.p, r.p);
  precondition(site(1, FIN_SET_RHO));
  SUCEM_LAUNCH((k_bcopy<NB>), G, TB, 0, c->stream, n, G, z.p, p.p);
  std::vector<int> done(NB, 0), broken(NB, 0);
  for (int j = 0; j < NB; ++j)
    if (j >= nb || bn[j] == 0.0) done[j] = 1;
  // device copy of the systems still iterating (the update kernels skip the others)
  DBuf<int> act(NB);
  auto upload_act = [&] {
    std::vector<int> a(NB);
    for (int j = 0; j < NB; ++j) a[j] = done[j] ? 0 : 1;
    h2d(c, act.p, a.data(), NB);
    c->sync();
  };
  upload_act();
  std::vector<double> rel(NB, 0.0);
  std::vector<int> iters(NB, 0);
  int it = 0;
  bool all_done = std::all_of(done.begin(), done.end(), [](int d) { return d != 0; });
  for (it = 0; it < o.max_iter && !all_done; ++it) {
    spmm(p.p, q.p, p.p, 1, site(2, FIN_COCG_ALPHA));
    SUCEM_LAUNCH((k_bxr<NB>), G, TB, 0, c->stream, n, G, scal.p, act.p, x.p, r.p, p.p, q.p);
    precondition(site(3, FIN_COCG_BETA, it));
    SUCEM_LAUNCH((k_bpz<NB>), G, TB, 0, c->stream, n, G, scal.p, act.p, z.p, p.p);
    if ((it + 1) % o.check_every != 0 && it + 1 != o.max_iter) continue;
    std::vector<double> h = read(hist.p + (int64_t)it * NB, NB);
    bool cand = true, froze = false;
    for (int j = 0; j < NB; ++j) {
      if (done[j]) continue;
      rel[j] = h[j] / bn[j];
      if (!std::isfinite(rel[j])) {
        // breakdown of this system's recursion: stop it, keep the others going
        done[j] = broken[j] = 1;
        iters[j] = it + 1;
        froze = true;
        continue;
      }
      if (rel[j] <= o.tol) {
        if (!iters[j]) iters[j] = it + 1;
      } else {
        cand = false;
      }
    }
    if (froze) upload_act();
    all_done = std::all_of(done.begin(), done.end(), [](int d) { return d != 0; });
    if (!cand || all_done) continue;
    // every recursion is below tol: confirm with the true residuals, restart the others
    spmm(x.p, ax.p, nullptr, 0, RedSite{});
    SUCEM_LAUNCH((k_bresid<NB>), G, TB, 0, c->stream, n, G, b.p, ax.p, mask, r.p,
                 site(0, FIN_NORM));
    finish(site(0, FIN_NORM));
    std::vector<double> tr = scal_slot(S_BNORM);
    all_done = true;
    for (int j = 0; j < NB; ++j) {
      if (j >= nb || bn[j] == 0.0 || broken[j]) continue;
      rel[j] = std::sqrt(tr[j]) / bn[j];
      if (rel[j] <= o.tol) done[j] = 1;
      else { all_done = false; iters[j] = 0; done[j] = 0; }
    }
    upload_act();
    if (!all_done) {
      precondition(site(1, FIN_SET_RHO));
      SUCEM_LAUNCH((k_bcopy<NB>), G, TB, 0, c->stream, n, G, z.p, p.p);
    }
  }
  SUCEM_LAUNCH((k_deinterleave<NB>), grid_for(ne, TB), TB, 0, c->stream, (int64_t)n, nb, x.p, X);
  c->sync();
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (int j = 0; j < nb; ++j) {
    st[j] = SolveStats();
    st[j].status = broken[j] ? -1 : (done[j] ? 0 : it);
    st[j].iterations = iters[j] ? iters[j] : it;
    st[j].rel_residual = rel[j];
    st[j].elapsed_ms = ms;
    st[j].spmv_hi = it;
  }
}

}  // namespace

void solve_system_batch(Context* c, System* sys, int64_t nb, const double* k0,
                        const double2* B, double2* X, const SolveOpts& o, SolveStats* st) {
  SUCEM_CHECK(sys->st64.p, SUCEM_ERR_INVALID, "S and T must be assembled");
  const int64_t n = sys->space->ndofs;
  if (o.precond != SUCEM_PRECOND_GRADIENT || o.precision != 64) {
    // any other solver: the systems one after another, each with the system's own
    // (matrix-free) operator -- with the auxiliary-space preconditioner 2-3x faster on a
    // 40^3 order-1 sweep than the interleaved gradient-preconditioned batch below
    for (int64_t j = 0; j < nb; ++j)
      solve_system(c, sys, driven_coef(k0[j]), B + j * n, X + j * n, o, st[j], nullptr, 0);
    return;
  }
  SUCEM_CHECK(o.method == SUCEM_METHOD_COCG, SUCEM_ERR_UNSUPPORTED,
              "interleaved batched solves run FP64 COCG with the gradient preconditioner");
  SUCEM_CHECK(o.max_iter > 0 && o.check_every > 0, SUCEM_ERR_INVALID, "bad iteration limits");
  for (int64_t off = 0; off < nb; off += 16) {
    const int m = (int)std::min<int64_t>(16, nb - off);
    const double2* Bc = B + off * n;
    double2* Xc = X + off * n;
    if (m > 8) solve_chunk<16>(c, sys, m, k0 + off, Bc, Xc, o, st + off);
    else if (m > 4) solve_chunk<8>(c, sys, m, k0 + off, Bc, Xc, o, st + off);
    else if (m > 2) solve_chunk<4>(c, sys, m, k0 + off, Bc, Xc, o, st + off);
    else if (m > 1) solve_chunk<2>(c, sys, m, k0 + off, Bc, Xc, o, st + off);
    else solve_chunk<1>(c, sys, m, k0 + off, Bc, Xc, o, st + off);
  }
}

}  // namespace sucem
