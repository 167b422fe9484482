// Krylov solvers behind LinalgSolvers (Utilities/LinalgSolvers.py:30-322) and the
// driven solve of examples/infinitesimal_dipole/driver.py:83.
//
//  * COCG (conjugate orthogonal CG) for complex-symmetric A (the driven A = S -
//    k0^2 T + i k0 S0 is complex symmetric): one SpMV per iteration.
//  * BiCGStab (the reference default, LinalgSolvers.py:239-248): two SpMVs.
//
// Precision: FP64 throughout, or FP32 vectors/matrix with FP64 reliable updates
// (x accumulated in FP64, true residual b - A x recomputed with the FP64 matrix
// whenever the FP32 recursion has reduced the residual by ru_delta, the FP32
// residual replaced by it and the recursion restarted from it).  All dot
// products accumulate in FP64.
//
// Every scalar (alpha, beta, ...) is derived on the device by each block from the
// per-block partial sums written by the previous kernel, so an iteration needs
// no host round trip; the host only polls the residual history every
// `check_every` iterations.
#include <chrono>
#include <cmath>

#include "internal.hpp"
#include "kernels.hpp"
#include "ops.hpp"

namespace sucem {
namespace {

constexpr int TB = 256;

__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  double den = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}
__device__ __forceinline__ double2 cmuld(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// rows [start, end) of logical block lb under the XCD-aware partition
struct Range { int start, end; };
__device__ __forceinline__ Range block_rows(int n, int G) {
  int lb = xcd_block(G);
  int chunk = (n + G - 1) / G;
  Range r;
  r.start = lb * chunk;
  r.end = min(n, r.start + chunk);
  return r;
}

template <typename C>
__device__ __forceinline__ C from_d(double2 v);
template <> __device__ __forceinline__ float2 from_d<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }
template <> __device__ __forceinline__ double2 from_d<double2>(double2 v) { return v; }

// ---- common vector kernels ------------------------------------------------------
template <typename C>
__global__ void k_zero_c(int64_t n, C* v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] = from_d<C>(make_double2(0, 0));
}

// part[lb] = (sum |b|^2 over unmasked rows, 0)
__global__ __launch_bounds__(TB) void k_norm2(int n, int G, const double2* b, const uint8_t* mask,
                                              double2* part) {
  Range rr = block_rows(n, G);
  double2 s = make_double2(0, 0);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB)
    if (!(mask && mask[i])) s.x += b[i].x * b[i].x + b[i].y * b[i].y;
  s = block_sum<TB>(s);
  if (threadIdx.x == 0) part[xcd_block(G)] = s;
}

// r = (b - Ax) (masked rows zero), stored in precision of C; rlo_dot partial = sum r (dinv r)
// (bilinear when !CONJ; conj(rt) r when CONJ with rt = r), true |r|^2 partial in FP64.
template <typename C>
__global__ __launch_bounds__(TB) void k_residual(int n, int G, const double2* b, const double2* ax,
                                                 const uint8_t* mask, const C* dinv, C* r,
                                                 double2* part_true, double2* part_rho, int cocg) {
  Range rr = block_rows(n, G);
  double2 tr = make_double2(0, 0), rho = make_double2(0, 0);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    double2 v = make_double2(b[i].x - ax[i].x, b[i].y - ax[i].y);
    if (mask && mask[i]) v = make_double2(0, 0);
    tr.x += v.x * v.x + v.y * v.y;
    C rl = from_d<C>(v);
    r[i] = rl;
    double2 rd = to_d(rl);
    if (cocg) {
      double2 z = to_d(cmul(dinv[i], rl));
      rho.x += rd.x * z.x - rd.y * z.y;
      rho.y += rd.x * z.y + rd.y * z.x;
    } else {
      rho.x += rd.x * rd.x + rd.y * rd.y;
    }
  }
  tr = block_sum<TB>(tr);
  rho = block_sum<TB>(rho);
  if (threadIdx.x == 0) {
    part_true[xcd_block(G)] = tr;
    if (part_rho) part_rho[xcd_block(G)] = rho;
  }
}

// x64 += xlo; xlo = 0  (FP32 reliable update)
template <typename C>
__global__ void k_accum(int n, int G, C* xlo, double2* x64) {
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    double2 a = to_d(xlo[i]);
    x64[i].x += a.x;
    x64[i].y += a.y;
    xlo[i] = from_d<C>(make_double2(0, 0));
  }
}

template <typename C>
__global__ void k_copy_cast(int n, int G, const C* src, C* dst) {
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) dst[i] = src[i];
}

template <typename C>
__global__ void k_store_x(int n, int G, const C* xlo, double2* x64) {
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) x64[i] = to_d(xlo[i]);
}

// ---- COCG -------------------------------------------------------------------
// p = dinv r (start)
template <typename C>
__global__ void k_cocg_start(int n, int G, const C* dinv, const C* r, C* p) {
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) p[i] = cmul(dinv[i], r[i]);
}

// alpha = rho / mu ; x += alpha p ; r -= alpha q ; partials of rho' = r.(dinv r), |r|^2
template <typename C>
__global__ __launch_bounds__(TB) void k_cocg_xr(int n, int G, const double2* prho_cur,
                                                const double2* pmu, C* x, C* r, const C* p,
                                                const C* q, const C* dinv, double2* prho_next,
                                                double2* prr_next) {
  const double2 rho = sum_partials<TB>(prho_cur, G);
  const double2 mu = sum_partials<TB>(pmu, G);
  const double2 al = cdiv(rho, mu);
  const C a = from_d<C>(al);
  Range rr = block_rows(n, G);
  double2 nr = make_double2(0, 0), n2 = make_double2(0, 0);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    C pi = p[i], qi = q[i], xi = x[i], ri = r[i];
    C ap = cmul(a, pi), aq = cmul(a, qi);
    xi.x += ap.x; xi.y += ap.y;
    ri.x -= aq.x; ri.y -= aq.y;
    x[i] = xi;
    r[i] = ri;
    double2 rd = to_d(ri), z = to_d(cmul(dinv[i], ri));
    nr.x += rd.x * z.x - rd.y * z.y;
    nr.y += rd.x * z.y + rd.y * z.x;
    n2.x += rd.x * rd.x + rd.y * rd.y;
  }
  nr = block_sum<TB>(nr);
  n2 = block_sum<TB>(n2);
  if (threadIdx.x == 0) {
    prho_next[xcd_block(G)] = nr;
    prr_next[xcd_block(G)] = n2;
  }
}

// beta = rho'/rho ; p = dinv r + beta p ; hist[it] = |r|
template <typename C>
__global__ __launch_bounds__(TB) void k_cocg_p(int n, int G, const double2* prho_cur,
                                               const double2* prho_next, const double2* prr_next,
                                               const C* dinv, const C* r, C* p, double* hist,
                                               int it) {
  const double2 rho = sum_partials<TB>(prho_cur, G);
  const double2 rn = sum_partials<TB>(prho_next, G);
  const double2 be = cdiv(rn, rho);
  if (blockIdx.x == 0) {
    const double2 r2 = sum_partials<TB>(prr_next, G);
    if (threadIdx.x == 0) hist[it] = sqrt(r2.x);
  }
  const C b = from_d<C>(be);
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    C z = cmul(dinv[i], r[i]), pi = p[i];
    C bp = cmul(b, pi);
    p[i] = mk(z.x + bp.x, z.y + bp.y);
  }
}

// ---- BiCGStab (preconditioned; SciPy's formulation) ----------------------------
// state[0..1] rho (parity), state[2] alpha, state[3] omega
template <typename C>
__global__ __launch_bounds__(TB) void k_bicg_p(int n, int G, int it, int restart,
                                               const double2* prho, double2* state, const C* r,
                                               C* p, const C* v, const C* dinv, C* phat) {
  const double2 rho = sum_partials<TB>(prho, G);
  double2 beta = make_double2(0, 0);
  if (!restart) {
    const double2 rho_prev = state[(it + 1) & 1];
    beta = cmuld(cdiv(rho, rho_prev), cdiv(state[2], state[3]));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) state[it & 1] = rho;
  const C b = from_d<C>(beta), om = from_d<C>(state[3]);
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    C ri = r[i];
    C pv;
    if (!restart) {
      C pi = p[i], ov = cmul(om, v[i]);
      C t = mk(pi.x - ov.x, pi.y - ov.y);
      C bt = cmul(b, t);
      pv = mk(ri.x + bt.x, ri.y + bt.y);
    } else {
      pv = ri;
    }
    p[i] = pv;
    phat[i] = cmul(dinv[i], pv);
  }
}

// alpha = rho / (rt^H v) ; s = r - alpha v ; shat = dinv s
template <typename C>
__global__ __launch_bounds__(TB) void k_bicg_s(int n, int G, int it, const double2* prho,
                                               const double2* pv, double2* state, const C* r,
                                               const C* v, const C* dinv, C* s, C* shat) {
  const double2 rho = sum_partials<TB>(prho, G);
  const double2 rv = sum_partials<TB>(pv, G);
  const double2 al = cdiv(rho, rv);
  if (blockIdx.x == 0 && threadIdx.x == 0) state[2] = al;
  const C a = from_d<C>(al);
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    C av = cmul(a, v[i]), ri = r[i];
    C si = mk(ri.x - av.x, ri.y - av.y);
    s[i] = si;
    shat[i] = cmul(dinv[i], si);
  }
}

// omega = (t^H s)/(t^H t) ; x += alpha phat + omega shat ; r = s - omega t ;
// partials rho' = rt^H r, |r|^2 ; hist
template <typename C>
__global__ __launch_bounds__(TB) void k_bicg_x(int n, int G, int it, const double2* pt,
                                               double2* state, C* x, C* r, const C* s,
                                               const C* t, const C* phat, const C* shat,
                                               const C* rt, double2* prho_next,
                                               double2* prr_next) {
  // pt[lb] = sum conj(s) t = conj(t^H s) ; pt[G + lb] = |t|^2
  const double2 d1 = sum_partials<TB>(pt, G);
  const double2 d2 = sum_partials<TB>(pt + G, G);
  const double2 om = make_double2(d1.x / d2.x, -d1.y / d2.x);
  if (blockIdx.x == 0 && threadIdx.x == 0) state[3] = om;
  const C o = from_d<C>(om), a = from_d<C>(state[2]);
  Range rr = block_rows(n, G);
  double2 nr = make_double2(0, 0), n2 = make_double2(0, 0);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    C ap = cmul(a, phat[i]), os = cmul(o, shat[i]), xi = x[i];
    xi.x += ap.x + os.x;
    xi.y += ap.y + os.y;
    x[i] = xi;
    C ot = cmul(o, t[i]), si = s[i];
    C ri = mk(si.x - ot.x, si.y - ot.y);
    r[i] = ri;
    double2 rd = to_d(ri), rtd = to_d(rt[i]);
    nr.x += rtd.x * rd.x + rtd.y * rd.y;   // conj(rt) r
    nr.y += rtd.x * rd.y - rtd.y * rd.x;
    n2.x += rd.x * rd.x + rd.y * rd.y;
  }
  nr = block_sum<TB>(nr);
  n2 = block_sum<TB>(n2);
  if (threadIdx.x == 0) {
    prho_next[xcd_block(G)] = nr;
    prr_next[xcd_block(G)] = n2;
  }
}

__global__ __launch_bounds__(TB) void k_hist(int G, const double2* prr, double* hist, int it) {
  const double2 r2 = sum_partials<TB>(prr, G);
  if (threadIdx.x == 0) hist[it] = sqrt(r2.x);
}

// ---- host driver ----------------------------------------------------------------
// Operator abstraction over the driven system and the generic CSR matrix.
struct Op {
  Context* c;
  int64_t n;
  int G, L;
  const uint8_t* mask;
  virtual void apply_lo(const void* x, void* y, const void* w, bool conj, double2* part) = 0;
  virtual void apply_hi(const double2* x, double2* y) = 0;
  virtual void dinv_lo(void* d) = 0;
  virtual ~Op() {}
};

template <typename R>
struct SysOp : Op {
  const System* sys;
  Coef k0;
  bool identity_prec = false;
  void apply_lo(const void* x, void* y, const void* w, bool conj, double2* part) override {
    sys_apply<R>(c, sys, k0, mask, x, y, w, conj, part, G, L);
  }
  void apply_hi(const double2* x, double2* y) override {
    sys_apply<double>(c, sys, k0, mask, x, y, nullptr, false, nullptr, G, L);
  }
  void dinv_lo(void* d) override {
    if (identity_prec) csr_dinv<R>(c, nullptr, mask, true, d, n);
    else sys_dinv<R>(c, sys, k0, mask, d);
  }
};

template <typename R>
struct CsrOp : Op {
  const Csr* A;
  bool identity_prec;
  void apply_lo(const void* x, void* y, const void* w, bool conj, double2* part) override {
    csr_apply<R>(c, A, mask, x, y, w, conj, part, G, L);
  }
  void apply_hi(const double2* x, double2* y) override {
    csr_apply<double>(c, A, mask, x, y, nullptr, false, nullptr, G, L);
  }
  void dinv_lo(void* d) override { csr_dinv<R>(c, A, mask, identity_prec, d, n); }
};

double sum_host(Context* c, const double2* part_d, int G, int comp = 0) {
  double* h = c->host_scratch(2 * G);
  SUCEM_HIP(hipMemcpyAsync(h, part_d, sizeof(double2) * G, hipMemcpyDeviceToHost, c->stream));
  c->sync();
  double s = 0;
  for (int i = 0; i < G; ++i) s += h[2 * i + comp];
  return s;
}

double read_one(Context* c, const double* d) {
  double* h = c->host_scratch(1);
  SUCEM_HIP(hipMemcpyAsync(h, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  c->sync();
  return h[0];
}

template <typename R>
void krylov(Op& op, const double2* b, double2* x, const SolveOpts& o, SolveStats& st,
            double* hist_h, int64_t hist_cap) {
  using C = typename C2<R>::T;
  Context* c = op.c;
  const int n = (int)op.n, G = op.G;
  const bool lo = sizeof(R) == 4;
  const bool cocg = o.method == SUCEM_METHOD_COCG;
  auto t0 = std::chrono::steady_clock::now();
  st = SolveStats();

  DBuf<double2> part(8 * G);               // scratch partials
  double2* pnorm = part.p;                 // G
  double2* ptrue = part.p + G;             // G
  double2* pmu = part.p + 2 * G;           // 2G (SpMV epilogue)
  double2* prho[2] = {part.p + 4 * G, part.p + 5 * G};
  double2* prr[2] = {part.p + 6 * G, part.p + 7 * G};
  DBuf<double> hist(o.max_iter + 1);
  DBuf<double2> state(4);
  DBuf<C> dinv(n), r(n), p(n), q(n);
  DBuf<C> xlo_buf(lo ? n : 0);
  DBuf<double2> ax(n);
  // BiCGStab extras
  DBuf<C> rt(cocg ? 0 : n), v(cocg ? 0 : n), phat(cocg ? 0 : n), s(cocg ? 0 : n),
      shat(cocg ? 0 : n), t(cocg ? 0 : n);
  C* xlo = lo ? xlo_buf.p : (C*)x;

  op.dinv_lo(dinv.p);
  SUCEM_LAUNCH((k_zero_c<double2>), grid_for(n, TB), TB, 0, c->stream, (int64_t)n, x);
  if (lo) SUCEM_LAUNCH((k_zero_c<C>), grid_for(n, TB), TB, 0, c->stream, (int64_t)n, xlo);
  SUCEM_LAUNCH(k_norm2, G, TB, 0, c->stream, n, G, b, op.mask, pnorm);
  const double bnorm = std::sqrt(sum_host(c, pnorm, G));
  if (bnorm == 0.0) {
    st.status = 0;
    st.rel_residual = 0;
    return;
  }
  // r = b - A x0 (x0 = 0)
  SUCEM_LAUNCH((k_zero_c<double2>), grid_for(n, TB), TB, 0, c->stream, (int64_t)n, ax.p);
  SUCEM_LAUNCH((k_residual<C>), G, TB, 0, c->stream, n, G, b, ax.p, op.mask, dinv.p, r.p, ptrue,
               prho[0], cocg ? 1 : 0);
  double last_true = 1.0;
  if (cocg) {
    SUCEM_LAUNCH((k_cocg_start<C>), G, TB, 0, c->stream, n, G, dinv.p, r.p, p.p);
  } else {
    SUCEM_LAUNCH((k_copy_cast<C>), G, TB, 0, c->stream, n, G, r.p, rt.p);
    // rho_0 = rt^H r = |r|^2 already in prho[0]
  }
  int it = 0;
  int restart = 1;
  st.status = o.max_iter;   // not converged unless we say otherwise
  double rel = 1.0;
  for (it = 0; it < o.max_iter; ++it) {
    if (cocg) {
      op.apply_lo(p.p, q.p, p.p, false, pmu);
      SUCEM_LAUNCH((k_cocg_xr<C>), G, TB, 0, c->stream, n, G, prho[it & 1], pmu, xlo, r.p, p.p, q.p,
                   dinv.p, prho[(it + 1) & 1], prr[(it + 1) & 1]);
      SUCEM_LAUNCH((k_cocg_p<C>), G, TB, 0, c->stream, n, G, prho[it & 1], prho[(it + 1) & 1],
                   prr[(it + 1) & 1], dinv.p, r.p, p.p, hist.p, it);
      st.spmv_lo += 1;
    } else {
      SUCEM_LAUNCH((k_bicg_p<C>), G, TB, 0, c->stream, n, G, it, restart, prho[it & 1], state.p,
                   r.p, p.p, v.p, dinv.p, phat.p);
      restart = 0;
      op.apply_lo(phat.p, v.p, rt.p, true, pmu);
      // pmu holds sum conj(rt) v
      SUCEM_LAUNCH((k_bicg_s<C>), G, TB, 0, c->stream, n, G, it, prho[it & 1], pmu, state.p, r.p,
                   v.p, dinv.p, s.p, shat.p);
      op.apply_lo(shat.p, t.p, s.p, true, pmu);
      SUCEM_LAUNCH((k_bicg_x<C>), G, TB, 0, c->stream, n, G, it, pmu, state.p, xlo, r.p, s.p, t.p,
                   phat.p, shat.p, rt.p, prho[(it + 1) & 1], prr[(it + 1) & 1]);
      SUCEM_LAUNCH(k_hist, 1, TB, 0, c->stream, G, prr[(it + 1) & 1], hist.p, it);
      st.spmv_lo += 2;
    }
    const bool check = ((it + 1) % o.check_every) == 0 || it + 1 == o.max_iter;
    if (!check) continue;
    rel = read_one(c, hist.p + it) / bnorm;
    if (!std::isfinite(rel)) {
      st.status = -1;   // breakdown
      break;
    }
    const bool ru = rel <= o.tol || (lo && rel <= o.ru_delta * last_true);
    if (!ru) continue;
    // reliable update / residual replacement with the FP64 operator
    if (lo) SUCEM_LAUNCH((k_accum<C>), G, TB, 0, c->stream, n, G, xlo, x);
    op.apply_hi(x, ax.p);
    st.spmv_hi += 1;
    st.reliable_updates += 1;
    SUCEM_LAUNCH((k_residual<C>), G, TB, 0, c->stream, n, G, b, ax.p, op.mask, dinv.p, r.p, ptrue,
                 cocg ? prho[(it + 1) & 1] : nullptr, cocg ? 1 : 0);
    const double true_rel = std::sqrt(sum_host(c, ptrue, G)) / bnorm;
    rel = true_rel;
    last_true = true_rel;
    if (true_rel <= o.tol) {
      st.status = 0;
      ++it;
      break;
    }
    if (cocg) {
      // restart the direction from the replaced residual: keeping the old direction
      // after a residual replacement stalls COCG on these indefinite systems
      SUCEM_LAUNCH((k_cocg_start<C>), G, TB, 0, c->stream, n, G, dinv.p, r.p, p.p);
    } else {
      // restart BiCGStab from the replaced residual (rt = r, rho = |r|^2)
      SUCEM_LAUNCH((k_residual<C>), G, TB, 0, c->stream, n, G, b, ax.p, op.mask, dinv.p, r.p, ptrue,
                   prho[(it + 1) & 1], 0);
      SUCEM_LAUNCH((k_copy_cast<C>), G, TB, 0, c->stream, n, G, r.p, rt.p);
      restart = 1;
    }
  }
  if (st.status != 0) {
    // final true residual for the report
    if (lo) SUCEM_LAUNCH((k_accum<C>), G, TB, 0, c->stream, n, G, xlo, x);
    op.apply_hi(x, ax.p);
    st.spmv_hi += 1;
    SUCEM_LAUNCH((k_residual<C>), G, TB, 0, c->stream, n, G, b, ax.p, op.mask, dinv.p, r.p, ptrue,
                 nullptr, 1);
    rel = std::sqrt(sum_host(c, ptrue, G)) / bnorm;
    if (st.status > 0 && rel <= o.tol) st.status = 0;
  }
  st.iterations = it;
  st.rel_residual = rel;
  if (hist_h && hist_cap > 0) {
    int64_t m = std::min<int64_t>(hist_cap, it);
    SUCEM_HIP(hipMemcpyAsync(hist_h, hist.p, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
  }
  c->sync();
  if (hist_h) {
    int64_t m = std::min<int64_t>(hist_cap, it);
    for (int64_t i = 0; i < m; ++i) hist_h[i] /= bnorm;
  }
  st.elapsed_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

void solve_system(Context* c, const System* sys, Coef k0, const double2* b, double2* x,
                  const SolveOpts& o, SolveStats& st, double* hist_h, int64_t hist_cap) {
  SUCEM_CHECK(sys->st64.p, SUCEM_ERR_INVALID, "S and T must be assembled");
  SUCEM_CHECK(o.method == SUCEM_METHOD_COCG || o.method == SUCEM_METHOD_BICGSTAB,
              SUCEM_ERR_UNSUPPORTED, "method must be COCG or BiCGStab");
  SUCEM_CHECK(o.precision == 32 || o.precision == 64, SUCEM_ERR_INVALID, "precision must be 32 or 64");
  SUCEM_CHECK(o.max_iter > 0 && o.check_every > 0, SUCEM_ERR_INVALID, "bad iteration limits");
  const Space* s = sys->space;
  auto fill = [&](Op& op) {
    op.c = c;
    op.n = s->ndofs;
    op.G = op_grid(c, s->ndofs);
    op.L = lanes_for(s->ndofs, s->nnz);
    op.mask = o.eliminate_pec ? s->pec.p : nullptr;
  };
  if (o.precision == 32) {
    SysOp<float> op;
    fill(op);
    op.sys = sys;
    op.k0 = k0;
    op.identity_prec = o.precond == SUCEM_PRECOND_NONE;
    krylov<float>(op, b, x, o, st, hist_h, hist_cap);
  } else {
    SysOp<double> op;
    fill(op);
    op.sys = sys;
    op.k0 = k0;
    op.identity_prec = o.precond == SUCEM_PRECOND_NONE;
    krylov<double>(op, b, x, o, st, hist_h, hist_cap);
  }
}

void solve_csr(Context* c, const Csr* A, const double2* b, double2* x, const SolveOpts& o,
               SolveStats& st, double* hist_h, int64_t hist_cap) {
  SUCEM_CHECK(o.method == SUCEM_METHOD_COCG || o.method == SUCEM_METHOD_BICGSTAB,
              SUCEM_ERR_UNSUPPORTED, "method must be COCG or BiCGStab");
  SUCEM_CHECK(o.precision == 32 || o.precision == 64, SUCEM_ERR_INVALID, "precision must be 32 or 64");
  SUCEM_CHECK(o.max_iter > 0 && o.check_every > 0, SUCEM_ERR_INVALID, "bad iteration limits");
  auto fill = [&](Op& op) {
    op.c = c;
    op.n = A->n;
    op.G = op_grid(c, A->n);
    op.L = lanes_for(A->n, A->nnz);
    op.mask = nullptr;
  };
  if (o.precision == 32) {
    SUCEM_CHECK(A->val32.p, SUCEM_ERR_INVALID, "matrix has no FP32 copy");
    CsrOp<float> op;
    fill(op);
    op.A = A;
    op.identity_prec = o.precond == SUCEM_PRECOND_NONE;
    krylov<float>(op, b, x, o, st, hist_h, hist_cap);
  } else {
    CsrOp<double> op;
    fill(op);
    op.A = A;
    op.identity_prec = o.precond == SUCEM_PRECOND_NONE;
    krylov<double>(op, b, x, o, st, hist_h, hist_cap);
  }
}

}  // namespace sucem
