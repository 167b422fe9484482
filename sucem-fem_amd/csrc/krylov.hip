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
// Scalars (alpha, beta, rho, omega) live on the device.  Every reducing kernel
// writes per-block partials; a one-block kernel launched behind it sums them in
// block order (deterministic) and updates the scalars, so an iteration is a fixed
// sequence of launches with no host round trip.  (Letting the last block of the
// producing kernel finish the sum needs a device-scope fence in every block, which
// writes back the XCD's L2 and cost up to half of an SpMV.)  The host polls the residual history every
// `check_every` iterations.
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>

#include "internal.hpp"
#include "kernels.hpp"
#include "ops.hpp"

namespace sucem {
namespace {

constexpr int TB = 256;

// a workspace slot viewed as an array of T
template <typename T>
struct Slot {
  T* p;
};

struct Range { int start, end; };
__device__ __forceinline__ Range block_rows(int n, int G) {
  int lb = xcd_block(G);
  int chunk = (n + G - 1) / G;
  Range r;
  r.start = lb * chunk;
  r.end = min(n, r.start + chunk);
  return r;
}

// ---- common vector kernels ------------------------------------------------------
template <typename C>
__global__ void k_zero_c(int64_t n, C* v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] = from_d<C>(make_double2(0, 0));
}

__global__ void k_zero_u32(int n, unsigned* p) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

__global__ void k_zero_f64(int n, double* p) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

// |b|^2 over unmasked rows
__global__ __launch_bounds__(TB) void k_norm2(int n, int G, const double2* b, const uint8_t* mask,
                                              RedSite rs) {
  Range rr = block_rows(n, G);
  double2 s = make_double2(0, 0);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB)
    if (!(mask && mask[i])) sq_acc(s, b[i], rs.paired);
  reduce_finalize<TB>(rs, G, s, make_double2(0, 0));
}

// r = b - Ax (masked rows zero), stored in the working precision.  Reduces the
// FP64 |r|^2 and rho of the stored residual: r.(dinv r) for COCG, |r|^2 otherwise.
template <typename C>
__global__ __launch_bounds__(TB) void k_residual(int n, int G, const double2* b, const double2* ax,
                                                 const uint8_t* mask, const C* dinv, C* r,
                                                 int cocg, RedSite rs) {
  Range rr = block_rows(n, G);
  double2 tr = make_double2(0, 0), rho = make_double2(0, 0);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    double2 v = make_double2(b[i].x - ax[i].x, b[i].y - ax[i].y);
    if (mask && mask[i]) v = make_double2(0, 0);
    sq_acc(tr, v, rs.paired);
    const C rl = from_d<C>(v);
    r[i] = rl;
    const double2 rd = to_d(rl);
    if (cocg) {
      // (paired systems have real diagonals: the complex product scales each part)
      const double2 z = to_d(cmul(dinv[i], rl));
      dot_acc(rho, rd, z, rs.paired);
    } else {
      rho.x += rd.x * rd.x + rd.y * rd.y;
    }
  }
  reduce_finalize<TB>(rs, G, tr, rho);
}

// x64 += sign * xlo (FP32 reliable update: the cycle's correction is added, and taken
// out again when the replacement shows it did not help); with clear, xlo = 0 afterwards
template <typename C>
__global__ void k_accum(int n, int G, C* xlo, double2* x64, double sign, int clear) {
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    const double2 a = to_d(xlo[i]);
    x64[i].x += sign * a.x;
    x64[i].y += sign * a.y;
    if (clear) xlo[i] = from_d<C>(make_double2(0, 0));
  }
}

template <typename C>
__global__ void k_copy(int n, int G, const C* src, C* dst) {
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) dst[i] = src[i];
}

// ---- COCG -------------------------------------------------------------------
// p = dinv r
template <typename C>
__global__ void k_cocg_start(int n, int G, const C* dinv, const C* r, C* p) {
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) p[i] = cmul(dinv[i], r[i]);
}

// x += alpha p ; r -= alpha q ; reduce rho' = r.(dinv r) and |r|^2 (-> beta)
template <typename C>
__global__ __launch_bounds__(TB) void k_cocg_xr(int n, int G, const double* scal, C* x, C* r,
                                                const C* p, const C* q, const C* dinv,
                                                RedSite rs) {
  const C a = from_d<C>(scal2(scal, S_ALPHA));
  Range rr = block_rows(n, G);
  double2 nr = make_double2(0, 0), n2 = make_double2(0, 0);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    const C pi = p[i], qi = q[i];
    C xi = x[i], ri = r[i];
    const C ap = cmul(a, pi), aq = cmul(a, qi);
    xi.x += ap.x; xi.y += ap.y;
    ri.x -= aq.x; ri.y -= aq.y;
    x[i] = xi;
    r[i] = ri;
    const double2 rd = to_d(ri), z = to_d(cmul(dinv[i], ri));
    nr.x += rd.x * z.x - rd.y * z.y;
    nr.y += rd.x * z.y + rd.y * z.x;
    n2.x += rd.x * rd.x + rd.y * rd.y;
  }
  reduce_finalize<TB>(rs, G, nr, n2);
}

// x += alpha p ; r -= alpha q  (the preconditioner kernel reduces afterwards)
template <typename C>
__global__ void k_cocg_xr_plain(int n, int G, const double* scal, C* x, C* r, const C* p,
                                const C* q) {
  const C a = from_d<C>(scal2(scal, S_ALPHA));
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    const C ap = cmul(a, p[i]), aq = cmul(a, q[i]);
    C xi = x[i], ri = r[i];
    xi.x += ap.x; xi.y += ap.y;
    ri.x -= aq.x; ri.y -= aq.y;
    x[i] = xi;
    r[i] = ri;
  }
}

// x += alpha p ; r -= alpha q after the patch operator left the sums of DOFs shared by
// several patches unsummed: for those rows (qcode >= 0) q is the sum of their partial sums,
// in patch order (zero on masked rows).  alpha = rho / (p.Ap) is formed by every block from
// the operator's Gd block partials (its p.Ap epilogue): all blocks sum them in one fixed
// order, so they agree bit for bit, and the one-block finalising launch between the two
// kernels is saved.  (Moving x += alpha p into the p update, which reads p anyway, was
// measured: this kernel 30.3 -> 23.4 us, the p update 12.2 -> 20.9 us per FP32 application
// at 40^3 order 2, so not kept.)
// rho is read from S_RHO_NEXT, which no block of this kernel writes; block 0 stores alpha
// and copies that rho to S_RHO, which the next p update reads as the previous rho.
template <typename C>
__global__ __launch_bounds__(TB) void k_cocg_xr_def_a(int n, int G, int Gd, RedSite rs, C* x,
                                                      C* r, const C* p, const C* q,
                                                      const int32_t* __restrict__ qcode,
                                                      const C* __restrict__ part,
                                                      const uint8_t* __restrict__ mask) {
  double2 s0 = make_double2(0, 0);
  for (int k = threadIdx.x; k < Gd; k += TB) {
    const double2 t = rs.part[k];
    s0.x += t.x;
    s0.y += t.y;
  }
  s0 = block_sum<TB>(s0);
  const double2 rho = scal2(rs.scal, S_RHO_NEXT);
  const bool pr = rs.paired;
  const double2 al = pr ? pdiv0(rho, s0) : cdiv0(rho, s0);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    rs.scal[S_ALPHA] = al.x;
    rs.scal[S_ALPHA + 1] = al.y;
    rs.scal[S_RHO] = rho.x;
    rs.scal[S_RHO + 1] = rho.y;
  }
  const C a = from_d<C>(al);
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    const int code = qcode[i];
    C qi;
    if (code < 0) {
      qi = q[i];
    } else {
      qi = mk(decltype(qi.x)(0), decltype(qi.x)(0));
      const int q0 = code >> 4, q1 = q0 + (code & 15);
      for (int j = q0; j < q1; ++j) {
        const C t = part[j];
        qi.x += t.x;
        qi.y += t.y;
      }
      if (mask && mask[i]) qi = mk(decltype(qi.x)(0), decltype(qi.x)(0));
    }
    const C pi = p[i];
    const C ap = pr ? pmul(a, pi) : cmul(a, pi), aq = pr ? pmul(a, qi) : cmul(a, qi);
    C xi = x[i], ri = r[i];
    xi.x += ap.x; xi.y += ap.y;
    ri.x -= aq.x; ri.y -= aq.y;
    x[i] = xi;
    r[i] = ri;
  }
}

// p = z + beta p with beta = (r.z) / rho formed by every block from the preconditioner's Gr
// block partials (r.z at part[0, Gr), |r|^2 at part[Gr, 2Gr)) in one fixed order; rho is read
// from S_RHO, which no block of this kernel writes.  Block 0 stores beta, the new rho (in
// S_RHO_NEXT, for the next alpha), |r|^2 and the residual history entry.
template <typename C>
__global__ __launch_bounds__(TB) void k_cocg_pz_b(int n, int G, int Gr, RedSite rs, const C* z,
                                                  C* p) {
  double2 s0 = make_double2(0, 0), s1 = make_double2(0, 0);
  for (int k = threadIdx.x; k < Gr; k += TB) {
    const double2 t = rs.part[k];
    s0.x += t.x;
    s0.y += t.y;
  }
  s0 = block_sum<TB>(s0);
  const bool pr = rs.paired;
  const double2 be = pr ? pdiv0(s0, scal2(rs.scal, S_RHO)) : cdiv0(s0, scal2(rs.scal, S_RHO));
  if (blockIdx.x == 0) {
    for (int k = threadIdx.x; k < Gr; k += TB) {
      const double2 t = rs.part[Gr + k];
      s1.x += t.x;
      s1.y += t.y;
    }
    s1 = block_sum<TB>(s1);
    if (threadIdx.x == 0) {
      double* sc = rs.scal;
      sc[S_BETA] = be.x; sc[S_BETA + 1] = be.y;
      sc[S_RHO_NEXT] = s0.x; sc[S_RHO_NEXT + 1] = s0.y;
      // paired: the larger of the two relative residuals, squared (see prel2)
      sc[S_RR] = pr ? prel2(sc, s1) : s1.x;
      if (rs.hist) rs.hist[rs.it] = sqrt(sc[S_RR]);
    }
  }
  const C b = from_d<C>(be);
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    const C pi = p[i], zi = z[i];
    const C bp = pr ? pmul(b, pi) : cmul(b, pi);
    p[i] = mk(zi.x + bp.x, zi.y + bp.y);
  }
}

// p = z + beta p
template <typename C>
__global__ void k_cocg_pz(int n, int G, const double* scal, const C* z, C* p) {
  const C b = from_d<C>(scal2(scal, S_BETA));
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    const C bp = cmul(b, p[i]), zi = z[i];
    p[i] = mk(zi.x + bp.x, zi.y + bp.y);
  }
}

// p = dinv r + beta p
template <typename C>
__global__ void k_cocg_p(int n, int G, const double* scal, const C* dinv, const C* r, C* p) {
  const C b = from_d<C>(scal2(scal, S_BETA));
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    const C z = cmul(dinv[i], r[i]);
    const C bp = cmul(b, p[i]);
    p[i] = mk(z.x + bp.x, z.y + bp.y);
  }
}

// ---- BiCGStab (preconditioned; SciPy's formulation) ----------------------------
// p = r + beta (p - omega v) (p = r on a restart) ; phat = dinv p
template <typename C>
__global__ void k_bicg_p(int n, int G, int restart, const double* scal, const C* r, C* p,
                         const C* v, const C* dinv, C* phat) {
  const C b = from_d<C>(scal2(scal, S_BETA)), om = from_d<C>(scal2(scal, S_OMEGA));
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    const C ri = r[i];
    C pv = ri;
    if (!restart) {
      const C pi = p[i], ov = cmul(om, v[i]);
      const C t = mk(pi.x - ov.x, pi.y - ov.y);
      const C bt = cmul(b, t);
      pv = mk(ri.x + bt.x, ri.y + bt.y);
    }
    p[i] = pv;
    phat[i] = cmul(dinv[i], pv);
  }
}

// s = r - alpha v ; shat = dinv s
template <typename C>
__global__ void k_bicg_s(int n, int G, const double* scal, const C* r, const C* v, const C* dinv,
                         C* s, C* shat) {
  const C a = from_d<C>(scal2(scal, S_ALPHA));
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    const C av = cmul(a, v[i]), ri = r[i];
    const C si = mk(ri.x - av.x, ri.y - av.y);
    s[i] = si;
    shat[i] = cmul(dinv[i], si);
  }
}

// x += alpha phat + omega shat ; r = s - omega t ; reduce conj(rt).r and |r|^2 (-> beta)
template <typename C>
__global__ __launch_bounds__(TB) void k_bicg_x(int n, int G, const double* scal, C* x, C* r,
                                               const C* s, const C* t, const C* phat,
                                               const C* shat, const C* rt, RedSite rs) {
  const C o = from_d<C>(scal2(scal, S_OMEGA)), a = from_d<C>(scal2(scal, S_ALPHA));
  Range rr = block_rows(n, G);
  double2 nr = make_double2(0, 0), n2 = make_double2(0, 0);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    const C ap = cmul(a, phat[i]), os = cmul(o, shat[i]);
    C xi = x[i];
    xi.x += ap.x + os.x;
    xi.y += ap.y + os.y;
    x[i] = xi;
    const C ot = cmul(o, t[i]), si = s[i];
    const C ri = mk(si.x - ot.x, si.y - ot.y);
    r[i] = ri;
    const double2 rd = to_d(ri), rtd = to_d(rt[i]);
    nr.x += rtd.x * rd.x + rtd.y * rd.y;   // conj(rt) r
    nr.y += rtd.x * rd.y - rtd.y * rd.x;
    n2.x += rd.x * rd.x + rd.y * rd.y;
  }
  reduce_finalize<TB>(rs, G, nr, n2);
}

// ---- host driver ----------------------------------------------------------------
// Operator abstraction over the driven system and the generic CSR matrix.
// One block of FIN_TB threads: the partials were written by other XCDs, so their
// loads go to memory; all of a thread's loads are issued before any is summed.
constexpr int FIN_TB = 1024;
__global__ __launch_bounds__(FIN_TB) void k_finalize(RedSite rs, int G) {
  double2 a = make_double2(0, 0), b = make_double2(0, 0);
  constexpr int K = 4;
  for (int i0 = threadIdx.x; i0 < G; i0 += K * FIN_TB) {
    double2 p[K], q[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = i0 + k * FIN_TB;
      p[k] = i < G ? rs.part[i] : make_double2(0, 0);
      q[k] = i < G ? rs.part[G + i] : make_double2(0, 0);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      a.x += p[k].x; a.y += p[k].y;
      b.x += q[k].x; b.y += q[k].y;
    }
  }
  a = block_sum<FIN_TB>(a);
  b = block_sum<FIN_TB>(b);
  if (threadIdx.x == 0) finalize(rs, a, b);
}

struct Op {
  Context* c = nullptr;
  int64_t n = 0;
  SpmvCfg flo, fhi;
  const uint8_t* mask = nullptr;
  virtual void apply_lo(const void* x, void* y, const void* w, int epi, const RedSite& rs) = 0;
  virtual void apply_hi(const double2* x, double2* y) = 0;
  virtual void dinv_lo(void* d) = 0;
  // gradient-space preconditioner (driven / shifted systems only)
  virtual int64_t nlag() { return 0; }
  virtual void winv_lo(void* winv) {}
  bool aux = false;   // the auxiliary-space form (weights of the gradient term differ)
  virtual void precond_lo(const void* winv, const void* dinv, const void* r, void* w, void* z,
                          int G, const RedSite& rs, const void* vinv, void* u) {}
  // vector nodal term of the auxiliary-space preconditioner (3 nv entries)
  virtual int64_t nnod() { return 0; }
  virtual void vinv_lo(void* vinv) {}
  // operator form that leaves shared rows as partial sums (COCG): block partials, or 0
  virtual int deferred_parts() { return 0; }
  virtual void apply_deferred(const void* x, void* y, const void* w, const RedSite& rs) {}
  virtual const int32_t* qcode() { return nullptr; }
  virtual const void* partials() { return nullptr; }
  virtual ~Op() {}
};

template <typename R>
struct SysOp : Op {
  const System* sys = nullptr;
  Coef k;
  Coef kp;   // the preconditioner's coefficients (k with the complex shift)
  bool identity_prec = false;
  void apply_lo(const void* x, void* y, const void* w, int epi, const RedSite& rs) override {
    sys_apply<R>(c, sys, k, mask, x, y, w, epi, rs, flo);
  }
  void apply_hi(const double2* x, double2* y) override {
    RedSite none{};
    sys_apply<double>(c, sys, k, mask, x, y, nullptr, 0, none, fhi);
  }
  void dinv_lo(void* d) override {
    if (identity_prec) csr_dinv<R>(c, nullptr, mask, true, d, n);
    else sys_dinv<R>(c, sys, kp, mask, d);
  }
  int64_t nlag() override {
    ensure_gradient(c, sys->space);
    ensure_entity_maps(c, sys->space);
    return sys->space->nlag;
  }
  void winv_lo(void* winv) override {
    grad_winv<R>(c, const_cast<System*>(sys), kp, mask != nullptr, aux, winv);
  }
  void precond_lo(const void* winv, const void* dinv, const void* r, void* w, void* z, int G,
                  const RedSite& rs, const void* vinv, void* u) override {
    grad_precond<R>(c, sys->space, winv, dinv, r, w, z, mask, G, rs, vinv, u);
  }
  int64_t nnod() override {
    ensure_nodal(c, sys->space);
    return 3 * sys->space->mesh->nv;
  }
  void vinv_lo(void* vinv) override {
    nodal_winv<R>(c, const_cast<System*>(sys), kp, mask != nullptr, vinv);
  }
  int deferred_parts() override { return pt_deferred_parts(sys, flo); }
  void apply_deferred(const void* x, void* y, const void* w, const RedSite& rs) override {
    pt_apply_deferred<R>(c, sys, k, mask, x, y, w, rs, flo);
  }
  const int32_t* qcode() override { return pt_qcode(sys, flo); }
  const void* partials() override { return pt_partials(sys); }
};

template <typename R>
struct CsrOp : Op {
  const Csr* A = nullptr;
  bool identity_prec = false;
  void apply_lo(const void* x, void* y, const void* w, int epi, const RedSite& rs) override {
    csr_apply<R>(c, A, mask, x, y, w, epi, rs, flo);
  }
  void apply_hi(const double2* x, double2* y) override {
    RedSite none{};
    csr_apply<double>(c, A, mask, x, y, nullptr, 0, none, fhi);
  }
  void dinv_lo(void* d) override { csr_dinv<R>(c, A, mask, identity_prec, d, n); }
};

double read_scalar(Context* c, const double* d) {
  double* h = c->host_scratch(1);
  SUCEM_HIP(hipMemcpyAsync(h, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  c->sync();
  return h[0];
}

// One pair of timing events, reused: start/stop bracket work on the stream, harvest reads
// the pair once the stream has been synchronised past it.
struct OpTimer {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool running = false, pending = false;
  double total_ms = 0;
  int64_t samples = 0;
  void start(Context* c) {
    if (pending || running) return;
    if (!e0) {
      SUCEM_HIP(hipEventCreate(&e0));
      SUCEM_HIP(hipEventCreate(&e1));
    }
    SUCEM_HIP(hipEventRecord(e0, c->stream));
    running = true;
  }
  void stop(Context* c) {
    if (!running) return;
    SUCEM_HIP(hipEventRecord(e1, c->stream));
    running = false;
    pending = true;
  }
  void harvest() {
    if (!pending) return;
    float ms = 0.f;
    SUCEM_HIP(hipEventElapsedTime(&ms, e0, e1));
    pending = false;
    total_ms += ms;
    ++samples;
  }
  ~OpTimer() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

enum Site { SITE_NORM = 0, SITE_RESID, SITE_SPMV1, SITE_SPMV2, SITE_UPDATE, SITE_COUNT };

template <typename R>
void krylov(Op& op, const double2* b, double2* x, const SolveOpts& o, SolveStats& st,
            double* hist_h, int64_t hist_cap) {
  using C = typename C2<R>::T;
  Context* c = op.c;
  const int n = (int)op.n;
  const int G = op_grid(c, n, 8);            // vector kernels
  const int Gm = std::max(G, op.flo.parts());  // partial buffer size
  const bool lo = sizeof(R) == 4;
  const bool cocg = o.method == SUCEM_METHOD_COCG;
  const bool aux = o.precond == SUCEM_PRECOND_AUXILIARY;
  const bool grad = o.precond == SUCEM_PRECOND_GRADIENT || aux;
  SUCEM_CHECK(!grad || cocg, SUCEM_ERR_UNSUPPORTED,
              "the gradient and auxiliary-space preconditioners are available with COCG");
  const int64_t nl = grad ? op.nlag() : 0;
  SUCEM_CHECK(!grad || nl > 0, SUCEM_ERR_UNSUPPORTED,
              "the gradient and auxiliary-space preconditioners need an assembled Nedelec "
              "system");
  const int64_t nn = aux ? op.nnod() : 0;
  // COCG with the patch operator: the shared rows' partial sums are added by the x/r update,
  // which also forms alpha from the operator's Gd block partials
  const int Gd = (cocg && grad) ? op.deferred_parts() : 0;
  const bool paired = o.paired != 0;
  SUCEM_CHECK(!paired || Gd > 0, SUCEM_ERR_UNSUPPORTED,
              "paired solves run COCG with the gradient or auxiliary-space preconditioner "
              "and the patch operator");
  auto t0 = std::chrono::steady_clock::now();
  st = SolveStats();
  const char* lg = std::getenv("SUCEM_LOG");   // 3: one line per residual replacement
  const int log_level = lg ? std::atoi(lg) : 0;

  // work vectors live in the context's workspace (slots 0..20): no device allocation (or
  // free) inside a solve once the context has run one of this size
  const Slot<double2> part{ws_take<double2>(c, 0, (int64_t)SITE_COUNT * 2 * Gm)};
  const Slot<unsigned> counters{ws_take<unsigned>(c, 1, SITE_COUNT)};
  const Slot<double> scal{ws_take<double>(c, 2, S_COUNT)};
  const Slot<double> hist{ws_take<double>(c, 3, o.max_iter + 1)};
  SUCEM_LAUNCH(k_zero_u32, 1, 64, 0, c->stream, (int)SITE_COUNT, counters.p);
  SUCEM_LAUNCH(k_zero_f64, 1, 64, 0, c->stream, (int)S_COUNT, scal.p);
  auto site = [&](int s_, int mode, int it = 0) {
    RedSite r;
    r.part = part.p + (int64_t)s_ * 2 * Gm;
    r.counter = counters.p + s_;
    r.scal = scal.p;
    r.hist = hist.p;
    r.mode = mode;
    r.it = it;
    r.paired = paired ? 1 : 0;
    return r;
  };

  auto vec = [&](int k, int64_t cnt) { return Slot<C>{ws_take<C>(c, k, cnt)}; };
  const Slot<C> dinv = vec(4, n), r = vec(5, n), p = vec(6, n), q = vec(7, n);
  const Slot<C> xlo_buf = vec(8, lo ? n : 0);
  const Slot<double2> ax{ws_take<double2>(c, 9, n)};
  const Slot<C> rt = vec(10, cocg ? 0 : n), v = vec(11, cocg ? 0 : n),
                phat = vec(12, cocg ? 0 : n), s = vec(13, cocg ? 0 : n),
                shat = vec(14, cocg ? 0 : n), t = vec(15, cocg ? 0 : n);
  C* xlo = lo ? xlo_buf.p : (C*)x;
  const Slot<C> z = vec(16, grad ? n : 0), w = vec(17, nl), winv = vec(18, nl),
                u = vec(19, nn), vinv = vec(20, nn);
  // z = M^-1 r, reducing r.z / |r|^2 into the site's scalars
  auto precondition = [&](const RedSite& rs) {
    op.precond_lo(winv.p, dinv.p, r.p, w.p, z.p, G, rs, aux ? vinv.p : nullptr, u.p);
  };

  op.dinv_lo(dinv.p);
  op.aux = aux;
  if (grad) op.winv_lo(winv.p);
  if (aux) op.vinv_lo(vinv.p);
  if (!o.warm) SUCEM_LAUNCH((k_zero_c<double2>), grid_for(n, TB), TB, 0, c->stream, (int64_t)n, x);
  if (lo) SUCEM_LAUNCH((k_zero_c<C>), grid_for(n, TB), TB, 0, c->stream, (int64_t)n, xlo);
  SUCEM_LAUNCH(k_norm2, G, TB, 0, c->stream, n, G, b, op.mask, site(SITE_NORM, FIN_NORM));
  finish_reduction(c, site(SITE_NORM, FIN_NORM), G);
  double bnorm = std::sqrt(read_scalar(c, scal.p + S_BNORM));
  if (paired) {
    // the kernels report the larger of the two systems' residuals, each relative to its
    // own |b| (prel2), so the relative figures below divide by 1
    const double b1 = std::sqrt(read_scalar(c, scal.p + S_BNORM + 1));
    bnorm = (bnorm == 0.0 && b1 == 0.0) ? 0.0 : 1.0;
  }
  if (bnorm == 0.0) {
    st.status = 0;
    st.rel_residual = 0;
    c->sync();
    return;
  }
  // r = b - A x0 (x0 = 0, or the caller's x with o.warm)
  if (o.warm) {
    op.apply_hi(x, ax.p);
    st.spmv_hi += 1;
  } else {
    SUCEM_LAUNCH((k_zero_c<double2>), grid_for(n, TB), TB, 0, c->stream, (int64_t)n, ax.p);
  }
  SUCEM_LAUNCH((k_residual<C>), G, TB, 0, c->stream, n, G, b, ax.p, op.mask, dinv.p, r.p,
               cocg ? 1 : 0, site(SITE_RESID, FIN_RESID));
  finish_reduction(c, site(SITE_RESID, FIN_RESID), G);
  const double start_rel =
      o.warm ? std::sqrt(read_scalar(c, scal.p + S_TRUE)) / bnorm : 1.0;
  if (start_rel <= o.tol) {
    st.status = 0;
    st.rel_residual = start_rel;
    c->sync();
    return;
  }
  if (grad) {
    precondition(site(SITE_UPDATE, FIN_SET_RHO));
    SUCEM_LAUNCH((k_copy<C>), G, TB, 0, c->stream, n, G, z.p, p.p);
  } else if (cocg) {
    SUCEM_LAUNCH((k_cocg_start<C>), G, TB, 0, c->stream, n, G, dinv.p, r.p, p.p);
  } else {
    SUCEM_LAUNCH((k_copy<C>), G, TB, 0, c->stream, n, G, r.p, rt.p);
  }
  double last_true = start_rel;
  // stagnation: the best true residual so far and the replacements since it last halved
  double best_true = start_rel;
  int flat = 0;
  int it = 0;
  int restart = 1;
  // FP32 refinement cycles: a cycle runs from one residual replacement to the next
  int cycle_start = 0, longest_cycle = 0;
  // best recursion residual of the current cycle and the iteration it was reached
  double cycle_best = 1e300;
  int cycle_best_it = 0;
  st.status = o.max_iter;
  double rel = 1.0;
  // Operator timing: at the first iteration of every check interval (from the second
  // interval on) the working-precision operator application (its kernels and its fused
  // reduction epilogue, not the one-block finalisation) is bracketed by two events on the
  // solve's stream; the pair is read back after the interval's host check has synchronised.
  OpTimer timer;
  auto apply_epi = [&](const void* xv, void* yv, const void* wv, int epi, const RedSite& rs) {
    const bool sample = it >= o.check_every && it % o.check_every == 0;
    if (sample) timer.start(c);
    op.apply_lo(xv, yv, wv, epi, rs);
    if (sample) timer.stop(c);
    finish_reduction(c, rs, op.flo.parts());
  };
  // grid of the preconditioner's z kernel in that loop (4 blocks per CU: 23.4 against
  // 24.7 us per FP32 application with the 8 of the other vector kernels, which lose more
  // at 4: the x/r update 34.3 against 28.5 us)
  const int Gz = op_grid(c, n, 4);
  for (it = 0; it < o.max_iter; ++it) {
    if (cocg && grad && Gd > 0) {
      const RedSite ra = site(SITE_SPMV1, FIN_COCG_ALPHA);
      const bool sample = it >= o.check_every && it % o.check_every == 0;
      if (sample) timer.start(c);
      op.apply_deferred(p.p, q.p, p.p, ra);
      if (sample) timer.stop(c);
      SUCEM_LAUNCH((k_cocg_xr_def_a<C>), G, TB, 0, c->stream, n, G, Gd, ra, xlo, r.p, p.p, q.p,
                   op.qcode(), (const C*)op.partials(), op.mask);
      const RedSite rb = site(SITE_UPDATE, FIN_DEFER, it);
      op.precond_lo(winv.p, dinv.p, r.p, w.p, z.p, Gz, rb, aux ? vinv.p : nullptr, u.p);
      SUCEM_LAUNCH((k_cocg_pz_b<C>), G, TB, 0, c->stream, n, G, Gz, rb, z.p, p.p);
      st.spmv_lo += 1;
    } else if (cocg && grad) {
      apply_epi(p.p, q.p, p.p, 1, site(SITE_SPMV1, FIN_COCG_ALPHA));
      SUCEM_LAUNCH((k_cocg_xr_plain<C>), G, TB, 0, c->stream, n, G, scal.p, xlo, r.p, p.p, q.p);
      precondition(site(SITE_UPDATE, FIN_COCG_BETA, it));
      SUCEM_LAUNCH((k_cocg_pz<C>), G, TB, 0, c->stream, n, G, scal.p, z.p, p.p);
      st.spmv_lo += 1;
    } else if (cocg) {
      apply_epi(p.p, q.p, p.p, 1, site(SITE_SPMV1, FIN_COCG_ALPHA));
      SUCEM_LAUNCH((k_cocg_xr<C>), G, TB, 0, c->stream, n, G, scal.p, xlo, r.p, p.p, q.p, dinv.p,
                   site(SITE_UPDATE, FIN_COCG_BETA, it));
      finish_reduction(c, site(SITE_UPDATE, FIN_COCG_BETA, it), G);
      SUCEM_LAUNCH((k_cocg_p<C>), G, TB, 0, c->stream, n, G, scal.p, dinv.p, r.p, p.p);
      st.spmv_lo += 1;
    } else {
      SUCEM_LAUNCH((k_bicg_p<C>), G, TB, 0, c->stream, n, G, restart, scal.p, r.p, p.p, v.p, dinv.p,
                   phat.p);
      restart = 0;
      apply_epi(phat.p, v.p, rt.p, 2, site(SITE_SPMV1, FIN_BICG_ALPHA));
      SUCEM_LAUNCH((k_bicg_s<C>), G, TB, 0, c->stream, n, G, scal.p, r.p, v.p, dinv.p, s.p, shat.p);
      apply_epi(shat.p, t.p, s.p, 2, site(SITE_SPMV2, FIN_BICG_OMEGA));
      SUCEM_LAUNCH((k_bicg_x<C>), G, TB, 0, c->stream, n, G, scal.p, xlo, r.p, s.p, t.p, phat.p,
                   shat.p, rt.p, site(SITE_UPDATE, FIN_BICG_BETA, it));
      finish_reduction(c, site(SITE_UPDATE, FIN_BICG_BETA, it), G);
      st.spmv_lo += 2;
    }
    const bool check = ((it + 1) % o.check_every) == 0 || it + 1 == o.max_iter;
    if (!check) continue;
    rel = read_scalar(c, hist.p + it) / bnorm;
    timer.harvest();
    if (!std::isfinite(rel)) {
      st.status = -1;   // breakdown
      break;
    }
    // An FP32 cycle whose recursion has not dropped by ru_delta within 3x the longest
    // completed cycle (at least 64 iterations) is cut short: its correction is folded in
    // by a replacement anyway and a fresh recursion starts from the true residual.  (Near
    // the FP64 floor of b - A x, or on the indefinite shifted cavity operator, an FP32
    // recursion can stagnate above its target for the rest of the iteration budget.)
    const int cyc = it + 1 - cycle_start;
    // the first cycle has no completed cycle to compare with: a fixed cap
    const int cap = longest_cycle > 0 ? std::max(64, 3 * longest_cycle)
                                      : std::max(1000, o.max_iter / 8);
    // An FP32 cycle whose recursion has not set a new low (by 10 %) for 256 iterations (or
    // twice the longest completed cycle) has stalled.  On the shifted cavity operator
    // S - sigma T of a fine mesh the preconditioned directions carry gradient components
    // about 1/(sigma h^2) larger than the rest, which S annihilates only to FP32 rounding:
    // the recursion then wanders instead of converging, and waiting for the fixed cap
    // would spend up to max_iter / 8 iterations on it.
    if (rel < 0.9 * cycle_best) {
      cycle_best = rel;
      cycle_best_it = it + 1;
    }
    const int stall = std::max(256, 2 * longest_cycle);
    const bool cut = lo && (cyc >= cap || it + 1 - cycle_best_it >= stall);
    // an FP32 recursion whose residual grew a thousandfold over the cycle's start has
    // diverged (COCG on an indefinite operator with an indefinite preconditioner has no
    // monotone residual; in FP32 it can run away): its correction is dropped and the
    // solve is finished in FP64 from the last accepted x
    if (lo && rel > 1e3 * last_true) {
      SUCEM_LAUNCH((k_zero_c<C>), grid_for(n, TB), TB, 0, c->stream, (int64_t)n, xlo);
      st.fp32_failed = 1;
      ++it;
      break;
    }
    const bool ru = rel <= o.tol || (lo && rel <= o.ru_delta * last_true) || cut;
    if (!ru) continue;
    if (!cut) longest_cycle = std::max(longest_cycle, cyc);
    cycle_start = it + 1;
    cycle_best = 1e300;
    cycle_best_it = it + 1;
    // reliable update / residual replacement with the FP64 operator
    if (lo) SUCEM_LAUNCH((k_accum<C>), G, TB, 0, c->stream, n, G, xlo, x, 1.0, 0);
    op.apply_hi(x, ax.p);
    st.spmv_hi += 1;
    st.reliable_updates += 1;
    SUCEM_LAUNCH((k_residual<C>), G, TB, 0, c->stream, n, G, b, ax.p, op.mask, dinv.p, r.p,
                 cocg ? 1 : 0, site(SITE_RESID, FIN_RESID));
    finish_reduction(c, site(SITE_RESID, FIN_RESID), G);
    const double true_rel = std::sqrt(read_scalar(c, scal.p + S_TRUE)) / bnorm;
    if (log_level >= 3)
      std::fprintf(stderr, "[sucem krylov] it %d: recursion %.3e, true %.3e%s (cycle %d)\n",
                   it + 1, rel, true_rel, cut ? " cut" : "", cyc);
    if (lo && !(true_rel < last_true)) {
      // an FP32 cycle that did not lower the true residual: take its correction out again
      // and finish in FP64 from the last accepted x
      SUCEM_LAUNCH((k_accum<C>), G, TB, 0, c->stream, n, G, xlo, x, -1.0, 1);
      st.fp32_failed = 1;
      ++it;
      break;
    }
    if (lo) SUCEM_LAUNCH((k_zero_c<C>), grid_for(n, TB), TB, 0, c->stream, (int64_t)n, xlo);
    rel = true_rel;
    last_true = true_rel;
    if (true_rel <= o.tol) {
      st.status = 0;
      ++it;
      break;
    }
    // The true residual has a floor set by the FP64 arithmetic of b - A x (it rises with
    // the operator's condition number).  A recursion that keeps reaching its target while
    // three replacements in a row fail to halve the best true residual is at that floor:
    // stop (status > 0, the iterations done) instead of running to max_iter.
    if (true_rel <= 0.5 * best_true) {
      best_true = true_rel;
      flat = 0;
    } else if (++flat >= 3) {
      if (lo) {
        st.fp32_failed = 1;   // FP32 cycles stopped paying: FP64 takes over
      } else {
        st.status = it + 1;
        st.stagnated = 1;
      }
      ++it;
      break;
    }
    // restart from the replaced residual (rho was recomputed by k_residual): keeping
    // the old direction after a residual replacement stalls these indefinite systems
    if (grad) {
      precondition(site(SITE_UPDATE, FIN_SET_RHO));
      SUCEM_LAUNCH((k_copy<C>), G, TB, 0, c->stream, n, G, z.p, p.p);
    } else if (cocg) {
      SUCEM_LAUNCH((k_cocg_start<C>), G, TB, 0, c->stream, n, G, dinv.p, r.p, p.p);
    } else {
      SUCEM_LAUNCH((k_copy<C>), G, TB, 0, c->stream, n, G, r.p, rt.p);
      restart = 1;
    }
  }
  if (st.fp32_failed) {
    // the caller finishes the solve in FP64 from x (the accepted corrections)
    st.iterations = it;
    c->sync();
    timer.harvest();
    st.op_samples = timer.samples;
    st.op_ms = timer.samples ? timer.total_ms / (double)timer.samples : 0.0;
    return;
  }
  if (st.status != 0) {
    // final true residual for the report
    if (lo) SUCEM_LAUNCH((k_accum<C>), G, TB, 0, c->stream, n, G, xlo, x, 1.0, 1);
    op.apply_hi(x, ax.p);
    st.spmv_hi += 1;
    SUCEM_LAUNCH((k_residual<C>), G, TB, 0, c->stream, n, G, b, ax.p, op.mask, dinv.p, r.p, 1,
                 site(SITE_RESID, FIN_RESID));
    finish_reduction(c, site(SITE_RESID, FIN_RESID), G);
    rel = std::sqrt(read_scalar(c, scal.p + S_TRUE)) / bnorm;
    if (st.status > 0 && rel <= o.tol) st.status = 0;
    if (it > o.max_iter) it = o.max_iter;
  }
  st.iterations = it;
  st.rel_residual = rel;
  if (hist_h && hist_cap > 0) {
    const int64_t m = std::min<int64_t>(hist_cap, it);
    SUCEM_HIP(hipMemcpyAsync(hist_h, hist.p, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    c->sync();
    for (int64_t i = 0; i < m; ++i) hist_h[i] /= bnorm;
  }
  c->sync();
  timer.harvest();
  st.op_samples = timer.samples;
  st.op_ms = timer.samples ? timer.total_ms / (double)timer.samples : 0.0;
  st.elapsed_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ---- GMRES (LinalgSolvers.GMRESSolver, LinalgSolvers.py:251-260) ----------------------
// Restarted GMRES(m) in FP64 with right preconditioning, so the quantity it minimises and
// monitors is the true residual ||b - A x||: Arnoldi with classical Gram-Schmidt applied
// twice (Hermitian inner products on the device, block partials summed on the host in
// block order), Givens rotations and the small triangular solve on the host, x updated by
// M^-1 (V y) at the end of each cycle and the residual recomputed from scratch before the
// next one.

// part[i G + lb] = sum_r conj(V_i[r]) w[r], i < m; basis vectors taken 8 at a time so w is
// read once per 8 of them
__global__ __launch_bounds__(TB) void k_zmultidot(int n, int G, int m, const double2* V,
                                                  const double2* w, double2* part) {
  Range rr = block_rows(n, G);
  for (int i0 = 0; i0 < m; i0 += 8) {
    const int cnt = min(8, m - i0);
    double2 acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = make_double2(0.0, 0.0);
    for (int r = rr.start + threadIdx.x; r < rr.end; r += TB) {
      const double2 wr = w[r];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k < cnt) {
          const double2 v = V[(int64_t)(i0 + k) * n + r];
          acc[k].x += v.x * wr.x + v.y * wr.y;
          acc[k].y += v.x * wr.y - v.y * wr.x;
        }
      }
    }
    for (int k = 0; k < cnt; ++k) {
      const double2 t = block_sum<TB>(acc[k]);
      if (threadIdx.x == 0) part[(int64_t)(i0 + k) * G + xcd_block(G)] = t;
    }
  }
}

// w -= sum_i h_i V_i (i < m)
__global__ void k_zsubtract(int n, int G, int m, const double2* h, const double2* V, double2* w) {
  Range rr = block_rows(n, G);
  for (int r = rr.start + threadIdx.x; r < rr.end; r += TB) {
    double2 a = w[r];
    for (int i = 0; i < m; ++i) {
      const double2 v = V[(int64_t)i * n + r], hi = h[i];
      a.x -= hi.x * v.x - hi.y * v.y;
      a.y -= hi.x * v.y + hi.y * v.x;
    }
    w[r] = a;
  }
}

// u = sum_i y_i V_i (i < m)
__global__ void k_zcombine(int n, int G, int m, const double2* y, const double2* V, double2* u) {
  Range rr = block_rows(n, G);
  for (int r = rr.start + threadIdx.x; r < rr.end; r += TB) {
    double2 a = make_double2(0.0, 0.0);
    for (int i = 0; i < m; ++i) {
      const double2 v = V[(int64_t)i * n + r], yi = y[i];
      a.x += yi.x * v.x - yi.y * v.y;
      a.y += yi.x * v.y + yi.y * v.x;
    }
    u[r] = a;
  }
}

__global__ void k_zscale(int n, int G, double s, const double2* in, double2* out) {
  Range rr = block_rows(n, G);
  for (int r = rr.start + threadIdx.x; r < rr.end; r += TB)
    out[r] = make_double2(s * in[r].x, s * in[r].y);
}

// r = b - ax (masked rows 0)
__global__ void k_zresid(int n, int G, const double2* b, const double2* ax, const uint8_t* mask,
                         double2* r) {
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    double2 v = make_double2(b[i].x - ax[i].x, b[i].y - ax[i].y);
    if (mask && mask[i]) v = make_double2(0.0, 0.0);
    r[i] = v;
  }
}

// out = dinv .* in (Jacobi), in place allowed
__global__ void k_zjacobi(int n, int G, const double2* dinv, const double2* in, double2* out) {
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) out[i] = cmul(dinv[i], in[i]);
}

__global__ void k_zadd(int n, int G, const double2* u, double2* x) {
  Range rr = block_rows(n, G);
  for (int i = rr.start + threadIdx.x; i < rr.end; i += TB) {
    x[i].x += u[i].x;
    x[i].y += u[i].y;
  }
}

void gmres(Op& op, const double2* b, double2* x, const SolveOpts& o, SolveStats& st,
           double* hist_h, int64_t hist_cap) {
  using cd = std::complex<double>;
  Context* c = op.c;
  const int n = (int)op.n;
  const int G = op_grid(c, n, 8);
  const int m = std::max(1, o.restart);
  const bool aux = o.precond == SUCEM_PRECOND_AUXILIARY;
  const bool grad = o.precond == SUCEM_PRECOND_GRADIENT || aux;
  const bool jac = o.precond == SUCEM_PRECOND_JACOBI;
  auto t0 = std::chrono::steady_clock::now();
  st = SolveStats();
  // workspace slots 21..31 (the Krylov solver's are 0..20)
  double2* V = ws_take<double2>(c, 21, (int64_t)(m + 1) * n);
  double2* w = ws_take<double2>(c, 22, n);
  double2* z = ws_take<double2>(c, 23, n);
  double2* ax = ws_take<double2>(c, 24, n);
  double2* dinv = ws_take<double2>(c, 25, (jac || grad) ? n : 0);
  double2* part = ws_take<double2>(c, 26, (int64_t)(m + 1) * G);
  double2* hd = ws_take<double2>(c, 27, m + 1);
  const int64_t nl = grad ? op.nlag() : 0;
  const int64_t nn = aux ? op.nnod() : 0;
  double2* wg = ws_take<double2>(c, 28, nl);
  double2* winv = ws_take<double2>(c, 29, nl);
  double2* ug = ws_take<double2>(c, 30, nn);
  double2* vinv = ws_take<double2>(c, 31, nn);
  SUCEM_CHECK(!grad || nl > 0, SUCEM_ERR_UNSUPPORTED,
              "the gradient and auxiliary-space preconditioners need an assembled Nedelec "
              "system");
  if (jac || grad) op.dinv_lo(dinv);
  op.aux = aux;
  if (grad) op.winv_lo(winv);
  if (aux) op.vinv_lo(vinv);
  // z = M^-1 v
  auto precondition = [&](const double2* v, double2* out) {
    if (grad) {
      RedSite none{};
      op.precond_lo(winv, dinv, v, wg, out, G, none, aux ? vinv : nullptr, ug);
    } else if (jac) {
      SUCEM_LAUNCH(k_zjacobi, G, TB, 0, c->stream, n, G, (const double2*)dinv, v, out);
    } else {
      SUCEM_HIP(hipMemcpyAsync(out, v, sizeof(double2) * n, hipMemcpyDeviceToDevice, c->stream));
    }
  };
  std::vector<double2> hp;
  auto dots = [&](const double2* base, int cnt, const double2* v) {
    SUCEM_LAUNCH(k_zmultidot, G, TB, 0, c->stream, n, G, cnt, base, v, part);
    hp.resize((size_t)cnt * G);
    SUCEM_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double2) * cnt * G, hipMemcpyDeviceToHost,
                             c->stream));
    c->sync();
    std::vector<cd> out(cnt, cd(0, 0));
    for (int i = 0; i < cnt; ++i)
      for (int g = 0; g < G; ++g) out[i] += cd(hp[(size_t)i * G + g].x, hp[(size_t)i * G + g].y);
    return out;
  };
  auto upload = [&](const std::vector<cd>& v) {
    std::vector<double2> t(v.size());
    for (size_t i = 0; i < v.size(); ++i) t[i] = make_double2(v[i].real(), v[i].imag());
    SUCEM_HIP(hipMemcpyAsync(hd, t.data(), sizeof(double2) * t.size(), hipMemcpyHostToDevice,
                             c->stream));
  };
  // ||b|| over the unmasked rows
  SUCEM_LAUNCH((k_zero_c<double2>), grid_for(n, TB), TB, 0, c->stream, (int64_t)n, ax);
  SUCEM_LAUNCH(k_zresid, G, TB, 0, c->stream, n, G, b, (const double2*)ax, op.mask, w);
  const double bnorm = std::sqrt(std::abs(dots(w, 1, w)[0]));
  if (!o.warm || bnorm == 0.0)
    SUCEM_LAUNCH((k_zero_c<double2>), grid_for(n, TB), TB, 0, c->stream, (int64_t)n, x);
  if (bnorm == 0.0) {
    st.status = 0;
    c->sync();
    return;
  }
  int it = 0;
  double rel = 1.0;
  st.status = o.max_iter;
  std::vector<cd> H((size_t)(m + 1) * m), g(m + 1), cs(m), sn(m);
  auto h = [&](int i, int j) -> cd& { return H[(size_t)i * m + j]; };
  while (true) {
    // true residual of the current x
    op.apply_hi(x, ax);
    st.spmv_hi += 1;
    SUCEM_LAUNCH(k_zresid, G, TB, 0, c->stream, n, G, b, (const double2*)ax, op.mask, w);
    const double beta = std::sqrt(std::abs(dots(w, 1, w)[0]));
    rel = beta / bnorm;
    if (!std::isfinite(rel)) { st.status = -1; break; }
    if (rel <= o.tol) { st.status = 0; break; }
    if (it >= o.max_iter) break;
    SUCEM_LAUNCH(k_zscale, G, TB, 0, c->stream, n, G, 1.0 / beta, (const double2*)w, V);
    std::fill(H.begin(), H.end(), cd(0, 0));
    std::fill(g.begin(), g.end(), cd(0, 0));
    g[0] = beta;
    int jend = 0;
    bool breakdown = false;
    for (int j = 0; j < m && it < o.max_iter; ++j) {
      double2* vj = V + (int64_t)j * n;
      precondition(vj, z);
      op.apply_hi(z, w);
      st.spmv_hi += 1;
      // classical Gram-Schmidt, twice
      for (int pass = 0; pass < 2; ++pass) {
        std::vector<cd> hh = dots(V, j + 1, w);
        for (int i = 0; i <= j; ++i) h(i, j) += hh[i];
        upload(hh);
        SUCEM_LAUNCH(k_zsubtract, G, TB, 0, c->stream, n, G, j + 1, (const double2*)hd,
                     (const double2*)V, w);
      }
      const double hn = std::sqrt(std::abs(dots(w, 1, w)[0]));
      h(j + 1, j) = hn;
      // rotations so far, then a new one zeroing h(j+1, j)
      for (int i = 0; i < j; ++i) {
        const cd a = h(i, j), bb = h(i + 1, j);
        h(i, j) = cs[i] * a + sn[i] * bb;
        h(i + 1, j) = -std::conj(sn[i]) * a + cs[i] * bb;
      }
      const cd a = h(j, j), bb = h(j + 1, j);
      const double r = std::sqrt(std::norm(a) + std::norm(bb));
      if (r == 0.0) {
        cs[j] = 1.0;
        sn[j] = 0.0;
      } else if (std::abs(a) == 0.0) {
        cs[j] = 0.0;
        sn[j] = std::conj(bb) / std::abs(bb);
      } else {
        cs[j] = std::abs(a) / r;
        sn[j] = (a / std::abs(a)) * std::conj(bb) / r;
      }
      h(j, j) = cs[j] * a + sn[j] * bb;
      h(j + 1, j) = 0.0;
      g[j + 1] = -std::conj(sn[j]) * g[j];
      g[j] = cs[j] * g[j];
      ++it;
      jend = j + 1;
      const double est = std::abs(g[j + 1]) / bnorm;
      if (hist_h && it - 1 < hist_cap) hist_h[it - 1] = est;
      if (hn == 0.0 || !std::isfinite(hn)) { breakdown = true; break; }
      if (est <= o.tol) break;
      SUCEM_LAUNCH(k_zscale, G, TB, 0, c->stream, n, G, 1.0 / hn, (const double2*)w,
                   V + (int64_t)(j + 1) * n);
    }
    // y = H^-1 g (upper triangular), x += M^-1 (V y)
    std::vector<cd> y(jend);
    for (int i = jend - 1; i >= 0; --i) {
      cd s = g[i];
      for (int k = i + 1; k < jend; ++k) s -= h(i, k) * y[k];
      y[i] = h(i, i) == cd(0, 0) ? cd(0, 0) : s / h(i, i);
    }
    upload(y);
    SUCEM_LAUNCH(k_zcombine, G, TB, 0, c->stream, n, G, jend, (const double2*)hd,
                 (const double2*)V, w);
    precondition(w, z);
    SUCEM_LAUNCH(k_zadd, G, TB, 0, c->stream, n, G, (const double2*)z, x);
    st.reliable_updates += 1;   // cycles
    (void)breakdown;
  }
  c->sync();
  st.iterations = it;
  st.rel_residual = rel;
  if (st.status > 0) st.status = it;
  st.elapsed_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void check_opts(const SolveOpts& o) {
  SUCEM_CHECK(o.method == SUCEM_METHOD_COCG || o.method == SUCEM_METHOD_BICGSTAB ||
                  o.method == SUCEM_METHOD_GMRES,
              SUCEM_ERR_UNSUPPORTED, "method must be COCG, BiCGStab or GMRES");
  SUCEM_CHECK(o.method != SUCEM_METHOD_GMRES || o.precision == 64, SUCEM_ERR_UNSUPPORTED,
              "GMRES runs in FP64 (precision 64)");
  SUCEM_CHECK(o.method != SUCEM_METHOD_GMRES || (o.restart >= 1 && o.restart <= 1000),
              SUCEM_ERR_INVALID, "GMRES restart length must be in [1, 1000]");
  SUCEM_CHECK(o.precision == 32 || o.precision == 64, SUCEM_ERR_INVALID,
              "precision must be 32 or 64");
  SUCEM_CHECK(o.max_iter > 0 && o.check_every > 0, SUCEM_ERR_INVALID, "bad iteration limits");
}

}  // namespace

void finish_reduction(Context* c, const RedSite& rs, int G) {
  if (rs.mode == FIN_NONE || rs.mode == FIN_DEFER) return;
  SUCEM_LAUNCH(k_finalize, 1, FIN_TB, 0, c->stream, rs, G);
}

void solve_system(Context* c, const System* sys, Coef k, const double2* b, double2* x,
                  const SolveOpts& o, SolveStats& st, double* hist_h, int64_t hist_cap) {
  SUCEM_CHECK(sys->st64.p, SUCEM_ERR_INVALID, "S and T must be assembled");
  check_opts(o);
  // the preconditioner's diagonals come from cs S + ct (1 - i beta) T + i c0 S0
  Coef kp = k;
  kp.ti = -std::max(0.0, o.precond_shift) * k.t;
  SUCEM_CHECK(!o.paired || (o.method == SUCEM_METHOD_COCG && k.s0 == 0.0 && kp.ti == 0.0),
              SUCEM_ERR_UNSUPPORTED,
              "paired solves run COCG on real coefficients (c0 = 0, no preconditioner shift)");
  const Space* s = sys->space;
  auto fill = [&](Op& op) {
    op.c = c;
    op.n = s->ndofs;
    op.mask = o.eliminate_pec ? s->pec.p : nullptr;
    op.fhi = sys_cfg<double>(c, sys);
  };
  if (o.method == SUCEM_METHOD_GMRES) {
    SysOp<double> op;
    fill(op);
    op.flo = op.fhi;
    op.sys = sys;
    op.k = k;
    op.kp = kp;
    op.identity_prec = o.precond == SUCEM_PRECOND_NONE;
    gmres(op, b, x, o, st, hist_h, hist_cap);
    return;
  }
  if (o.precision == 32) {
    SysOp<float> op;
    fill(op);
    op.flo = sys_cfg<float>(c, sys);
    op.sys = sys;
    op.k = k;
    op.kp = kp;
    op.identity_prec = o.precond == SUCEM_PRECOND_NONE;
    krylov<float>(op, b, x, o, st, hist_h, hist_cap);
    if (st.fp32_failed) {
      // finish in FP64, warm-started from the FP32 cycles' accepted solution
      SolveOpts o64 = o;
      o64.precision = 64;
      o64.warm = 1;
      o64.max_iter = std::max(1, o.max_iter - st.iterations);
      SysOp<double> op64;
      fill(op64);
      op64.flo = op64.fhi;
      op64.sys = sys;
      op64.k = k;
      op64.kp = kp;
      op64.identity_prec = op.identity_prec;
      SolveStats s64;
      krylov<double>(op64, b, x, o64, s64, nullptr, 0);
      const SolveStats s32 = st;
      st = s64;
      st.iterations += s32.iterations;
      st.reliable_updates += s32.reliable_updates;
      st.spmv_lo += s32.spmv_lo;
      st.spmv_hi += s32.spmv_hi;
      st.elapsed_ms += s32.elapsed_ms;
      st.fp32_failed = 1;
      st.op_ms = s32.op_ms;
      st.op_samples = s32.op_samples;
      if (st.status > 0) st.status = st.iterations;
    }
  } else {
    SysOp<double> op;
    fill(op);
    op.flo = op.fhi;
    op.sys = sys;
    op.k = k;
    op.kp = kp;
    op.identity_prec = o.precond == SUCEM_PRECOND_NONE;
    krylov<double>(op, b, x, o, st, hist_h, hist_cap);
  }
}

// Paired solves run the patch operator's COCG form: in the working precision and, for FP32
// solves, in the FP64 finish too.
bool paired_solves_available(Context* c, const System* sys, int precision) {
  if (sys->space->order != 1 && sys->space->order != 2) return false;
  const bool p64 = pt_deferred_parts(sys, sys_cfg<double>(c, sys)) > 0;
  if (precision == 64) return p64;
  return p64 && pt_deferred_parts(sys, sys_cfg<float>(c, sys)) > 0;
}

// z = M^-1 r with the auxiliary-space preconditioner of cs S + ct T + i c0 S0 (diagonals
// built with the complex shift `shift`), in the given precision: the kernels COCG runs,
// applied once (r and z complex64 for precision 32, complex128 for 64).
void precond_apply(Context* c, const System* sys, Coef k, double shift, int precision,
                   bool eliminate_pec, const void* r, void* z) {
  SUCEM_CHECK(sys->st64.p, SUCEM_ERR_INVALID, "S and T must be assembled");
  SUCEM_CHECK(precision == 32 || precision == 64, SUCEM_ERR_INVALID,
              "precision must be 32 or 64");
  SUCEM_CHECK(shift >= 0, SUCEM_ERR_INVALID, "shift must be >= 0");
  Coef kp = k;
  kp.ti = -shift * k.t;
  auto run = [&](auto tag) {
    using R = decltype(tag);
    using C = typename C2<R>::T;
    SysOp<R> op;
    op.c = c;
    op.n = sys->space->ndofs;
    op.mask = eliminate_pec ? sys->space->pec.p : nullptr;
    op.sys = sys;
    op.k = k;
    op.kp = kp;
    op.aux = true;
    const int64_t n = op.n, nl = op.nlag(), nn = op.nnod();
    DBuf<C> dinv(n), w(nl), winv(nl), u(nn), vinv(nn);
    op.dinv_lo(dinv.p);
    op.winv_lo(winv.p);
    op.vinv_lo(vinv.p);
    RedSite none{};
    op.precond_lo(winv.p, dinv.p, r, w.p, z, op_grid(c, n, 8), none, vinv.p, u.p);
    c->sync();
  };
  if (precision == 32) run(float{});
  else run(double{});
}

void solve_csr(Context* c, const Csr* A, const double2* b, double2* x, const SolveOpts& o,
               SolveStats& st, double* hist_h, int64_t hist_cap) {
  check_opts(o);
  SUCEM_CHECK(!o.paired, SUCEM_ERR_UNSUPPORTED, "paired solves run on assembled systems");
  auto fill = [&](Op& op) {
    op.c = c;
    op.n = A->n;
    op.mask = nullptr;
    op.flo = op.fhi = csr_cfg(c, A);
  };
  if (o.method == SUCEM_METHOD_GMRES) {
    SUCEM_CHECK(o.precond == SUCEM_PRECOND_NONE || o.precond == SUCEM_PRECOND_JACOBI,
                SUCEM_ERR_UNSUPPORTED, "GMRES on a generic matrix: no or Jacobi preconditioning");
    CsrOp<double> op;
    fill(op);
    op.A = A;
    op.identity_prec = o.precond == SUCEM_PRECOND_NONE;
    gmres(op, b, x, o, st, hist_h, hist_cap);
    return;
  }
  if (o.precision == 32) {
    SUCEM_CHECK(A->val32.p, SUCEM_ERR_INVALID, "matrix has no FP32 copy");
    CsrOp<float> op;
    fill(op);
    op.A = A;
    op.identity_prec = o.precond == SUCEM_PRECOND_NONE;
    krylov<float>(op, b, x, o, st, hist_h, hist_cap);
    if (st.fp32_failed) {
      SolveOpts o64 = o;
      o64.precision = 64;
      o64.warm = 1;
      o64.max_iter = std::max(1, o.max_iter - st.iterations);
      CsrOp<double> op64;
      fill(op64);
      op64.A = A;
      op64.identity_prec = op.identity_prec;
      SolveStats s64;
      krylov<double>(op64, b, x, o64, s64, nullptr, 0);
      const SolveStats s32 = st;
      st = s64;
      st.iterations += s32.iterations;
      st.reliable_updates += s32.reliable_updates;
      st.spmv_lo += s32.spmv_lo;
      st.spmv_hi += s32.spmv_hi;
      st.elapsed_ms += s32.elapsed_ms;
      st.fp32_failed = 1;
      if (st.status > 0) st.status = st.iterations;
    }
  } else {
    CsrOp<double> op;
    fill(op);
    op.A = A;
    op.identity_prec = o.precond == SUCEM_PRECOND_NONE;
    krylov<double>(op, b, x, o, st, hist_h, hist_cap);
  }
}

}  // namespace sucem
