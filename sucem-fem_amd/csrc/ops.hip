// Sparse operators, point location and point-source right-hand sides.
//
// The driven operator is never materialised: y = (S - k0^2 T) x + i k0 S0 x
// streams the interleaved real (S, T) values of the shared pattern once and adds
// the compact ABC rows, replacing the three CSR conversions and two sparse adds
// of DrivenProblemABC.get_LHS_matrix (ProblemConfigurations/EMDrivenProblem.py:67-72).
#include <cmath>
#include <mutex>

#include "internal.hpp"
#include "kernels.hpp"
#include "ops.hpp"

namespace sucem {
namespace {

constexpr int TB = 256;

template <typename R> struct ST2;
template <> struct ST2<float> { using T = float2; };
template <> struct ST2<double> { using T = double2; };

// One group of L lanes per row, U loads per lane issued before any is used;
// each XCD gets a contiguous block of rows.  Epilogue (EPI > 0): per-row
// op(w_r) * y_r (op = conj when EPI == 2) and |y_r|^2 as per-block partials; a
// one-block kernel launched behind it finishes the sum and updates the Krylov scalars.
template <typename R, int L, int U, int EPI>
__global__ __launch_bounds__(TB) void k_sys_spmv(
    int n, int G, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
    const typename ST2<R>::T* __restrict__ st, const int32_t* __restrict__ s0p,
    const int32_t* __restrict__ s0c, const R* __restrict__ s0v, R cs, R ct, R c0,
    const uint8_t* __restrict__ mask, const typename C2<R>::T* __restrict__ x,
    typename C2<R>::T* __restrict__ y, const typename C2<R>::T* __restrict__ w, RedSite rs) {
  using C = typename C2<R>::T;
  using V = typename ST2<R>::T;
  constexpr int RPB = TB / L;
  const int lb = xcd_block(G);
  int chunk = (n + G - 1) / G;
  chunk = (chunk + RPB - 1) / RPB * RPB;
  const int start = lb * chunk;
  const int end = min(n, start + chunk);
  const int lane = threadIdx.x % L;
  double2 d1 = make_double2(0.0, 0.0), d2 = make_double2(0.0, 0.0);
  for (int base = start; base < end; base += RPB) {
    const int r = base + threadIdx.x / L;
    C as = mk(R(0), R(0)), at = mk(R(0), R(0)), a0 = mk(R(0), R(0));
    if (r < end) {
      const int j1 = rp[r + 1];
      for (int jb = rp[r] + lane; jb < j1; jb += L * U) {
        int cc[U];
        V vv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int j = jb + u * L;
          const bool ok = j < j1;
          cc[u] = ok ? col[j] : -1;
          vv[u] = ok ? st[j] : V{R(0), R(0)};
        }
        C xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = cc[u] >= 0 ? x[cc[u]] : mk(R(0), R(0));
#pragma unroll
        for (int u = 0; u < U; ++u) {
          as.x += vv[u].x * xv[u].x; as.y += vv[u].x * xv[u].y;
          at.x += vv[u].y * xv[u].x; at.y += vv[u].y * xv[u].y;
        }
      }
      if (s0p) {
        const int k1 = s0p[r + 1];
        for (int j = s0p[r] + lane; j < k1; j += L) {
          const R v = s0v[j];
          const C xc = x[s0c[j]];
          a0.x += v * xc.x; a0.y += v * xc.y;
        }
      }
    }
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) {
      as.x += __shfl_xor(as.x, o); as.y += __shfl_xor(as.y, o);
      at.x += __shfl_xor(at.x, o); at.y += __shfl_xor(at.y, o);
      a0.x += __shfl_xor(a0.x, o); a0.y += __shfl_xor(a0.y, o);
    }
    if (lane == 0 && r < end) {
      C yr = mk(cs * as.x + ct * at.x - c0 * a0.y, cs * as.y + ct * at.y + c0 * a0.x);
      if (mask && mask[r]) yr = mk(R(0), R(0));
      y[r] = yr;
      if (EPI) {
        const double2 yd = to_d(yr);
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

// Generic complex CSR: y = A x with the same structure and epilogue.
template <typename R, int L, int U, int EPI>
__global__ __launch_bounds__(TB) void k_csr_spmv(int n, int G, const int32_t* __restrict__ rp,
                                                 const int32_t* __restrict__ col,
                                                 const typename C2<R>::T* __restrict__ val,
                                                 const uint8_t* __restrict__ mask,
                                                 const typename C2<R>::T* __restrict__ x,
                                                 typename C2<R>::T* __restrict__ y,
                                                 const typename C2<R>::T* __restrict__ w,
                                                 RedSite rs) {
  using C = typename C2<R>::T;
  constexpr int RPB = TB / L;
  const int lb = xcd_block(G);
  int chunk = (n + G - 1) / G;
  chunk = (chunk + RPB - 1) / RPB * RPB;
  const int start = lb * chunk;
  const int end = min(n, start + chunk);
  const int lane = threadIdx.x % L;
  double2 d1 = make_double2(0.0, 0.0), d2 = make_double2(0.0, 0.0);
  for (int base = start; base < end; base += RPB) {
    const int r = base + threadIdx.x / L;
    C acc = mk(R(0), R(0));
    if (r < end) {
      const int j1 = rp[r + 1];
      for (int jb = rp[r] + lane; jb < j1; jb += L * U) {
        int cc[U];
        C aa[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int j = jb + u * L;
          const bool ok = j < j1;
          cc[u] = ok ? col[j] : -1;
          aa[u] = ok ? val[j] : mk(R(0), R(0));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const C xc = cc[u] >= 0 ? x[cc[u]] : mk(R(0), R(0));
          acc.x += aa[u].x * xc.x - aa[u].y * xc.y;
          acc.y += aa[u].x * xc.y + aa[u].y * xc.x;
        }
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

template <typename R>
__global__ void k_sys_dinv(int64_t n, const int32_t* diag, const typename ST2<R>::T* st,
                           const int32_t* s0diag, const R* s0v, double cs, double ct,
                           double cti, double c0, const uint8_t* mask,
                           typename C2<R>::T* dinv) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    const auto v = st[diag[r]];
    double re = cs * (double)v.x + ct * (double)v.y;
    double im = cti * (double)v.y;
    if (s0diag && s0diag[r] >= 0) im += c0 * (double)s0v[s0diag[r]];
    double den = re * re + im * im;
    bool off = (mask && mask[r]) || den == 0.0;
    dinv[r] = off ? mk(R(0), R(0)) : mk(R(re / den), R(-im / den));
  }
}

template <typename R>
__global__ void k_csr_dinv(int64_t n, const int32_t* diag, const typename C2<R>::T* val,
                           const uint8_t* mask, typename C2<R>::T* dinv) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    double re = 0, im = 0;
    if (diag[r] >= 0) { re = val[diag[r]].x; im = val[diag[r]].y; }
    double den = re * re + im * im;
    bool off = (mask && mask[r]) || den == 0.0;
    dinv[r] = off ? mk(R(0), R(0)) : mk(R(re / den), R(-im / den));
  }
}

template <typename R>
__global__ void k_csr_identity_dinv(int64_t n, const uint8_t* mask, typename C2<R>::T* dinv) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x)
    dinv[r] = (mask && mask[r]) ? mk(R(0), R(0)) : mk(R(1), R(0));
}

__global__ void k_locate(int64_t nc, const double* xyz, const int32_t* cells, int64_t n,
                         const double* pts, double tol, int32_t* out) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int32_t* cell = cells + 4 * c;
    const double* x0 = xyz + 3 * (int64_t)cell[0];
    double J[3][3];
    for (int k = 0; k < 3; ++k) {
      const double* xk = xyz + 3 * (int64_t)cell[k + 1];
      for (int r = 0; r < 3; ++r) J[r][k] = xk[r] - x0[r];
    }
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    double id = 1.0 / det;
    double Ji[3][3] = {{c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id,
                        (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id},
                       {c01 * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id,
                        (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
                       {c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id,
                        (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
    for (int64_t p = 0; p < n; ++p) {
      double d[3] = {pts[3 * p] - x0[0], pts[3 * p + 1] - x0[1], pts[3 * p + 2] - x0[2]};
      double l1 = Ji[0][0] * d[0] + Ji[0][1] * d[1] + Ji[0][2] * d[2];
      double l2 = Ji[1][0] * d[0] + Ji[1][1] * d[1] + Ji[1][2] * d[2];
      double l3 = Ji[2][0] * d[0] + Ji[2][1] * d[1] + Ji[2][2] * d[2];
      double l0 = 1.0 - l1 - l2 - l3;
      if (l0 >= -tol && l1 >= -tol && l2 >= -tol && l3 >= -tol) atomicMin(&out[p], (int32_t)c);
    }
  }
}

// Physical basis values N_i(x) = J^-T N^_i(x^) of the cell containing x.
__device__ void basis_values(const double* xyz, const int32_t* cell, const double* ref_basis,
                             int nloc, const double* p, double out[20][3]) {
  const double* x0 = xyz + 3 * (int64_t)cell[0];
  double J[3][3];
  for (int k = 0; k < 3; ++k) {
    const double* xk = xyz + 3 * (int64_t)cell[k + 1];
    for (int r = 0; r < 3; ++r) J[r][k] = xk[r] - x0[r];
  }
  double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  double id = 1.0 / det;
  double Ji[3][3] = {{c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id,
                      (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id},
                     {c01 * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id,
                      (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
                     {c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id,
                      (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
  double d[3] = {p[0] - x0[0], p[1] - x0[1], p[2] - x0[2]};
  double xh[3];
  for (int a = 0; a < 3; ++a) xh[a] = Ji[a][0] * d[0] + Ji[a][1] * d[1] + Ji[a][2] * d[2];
  const double mono[10] = {1.0,           xh[0],         xh[1],         xh[2],
                           xh[0] * xh[0], xh[0] * xh[1], xh[0] * xh[2], xh[1] * xh[1],
                           xh[1] * xh[2], xh[2] * xh[2]};
  for (int i = 0; i < nloc; ++i) {
    double nh[3];
    for (int dd = 0; dd < 3; ++dd) {
      double s = 0;
      for (int m = 0; m < 10; ++m) s += ref_basis[(i * 3 + dd) * 10 + m] * mono[m];
      nh[dd] = s;
    }
    for (int a = 0; a < 3; ++a) out[i][a] = Ji[0][a] * nh[0] + Ji[1][a] * nh[1] + Ji[2][a] * nh[2];
  }
}

__global__ void k_basis_at(int64_t n, const double* pts, const int32_t* cellid, const double* xyz,
                           const int32_t* cells, const double* ref_basis, int nloc, double* vals) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    double v[20][3];
    basis_values(xyz, cells + 4 * (int64_t)cellid[p], ref_basis, nloc, pts + 3 * p, v);
    for (int i = 0; i < nloc; ++i)
      for (int a = 0; a < 3; ++a) vals[(p * nloc + i) * 3 + a] = v[i][a];
  }
}

// One wavefront per right-hand side; its sources are added in input order
// (b[dof] += scale * N_dof(x_s) . value_s), so accumulation is deterministic.
__global__ __launch_bounds__(64) void k_rhs(int64_t nsrc, const double* pos, const double2* val,
                                            const int32_t* rhs, const int32_t* cellid,
                                            int64_t nrhs, double2 scale, const double* xyz,
                                            const int32_t* cells, const int32_t* cell_dofs,
                                            const double* ref_basis, int nloc, int64_t ndofs,
                                            double2* b) {
  for (int64_t r = blockIdx.x; r < nrhs; r += gridDim.x) {
    for (int64_t s = 0; s < nsrc; ++s) {
      if (rhs[s] != r) continue;
      const int32_t c = cellid[s];
      if (threadIdx.x < nloc) {
        double v[20][3];
        basis_values(xyz, cells + 4 * (int64_t)c, ref_basis, nloc, pos + 3 * s, v);
        const int i = threadIdx.x;
        double2 acc = make_double2(0, 0);
        for (int a = 0; a < 3; ++a) {
          acc.x += v[i][a] * val[3 * s + a].x;
          acc.y += v[i][a] * val[3 * s + a].y;
        }
        const int64_t dof = cell_dofs[(int64_t)c * nloc + i];
        double2 add = make_double2(scale.x * acc.x - scale.y * acc.y,
                                   scale.x * acc.y + scale.y * acc.x);
        double2 old = b[r * ndofs + dof];
        b[r * ndofs + dof] = make_double2(old.x + add.x, old.y + add.y);
      }
      __syncthreads();
    }
  }
}

template <typename R, int L, int U>
void launch_sys(Context* c, const System* sys, Coef k, const uint8_t* mask, const void* x,
                void* y, const void* w, int epi, const RedSite& rs, int G) {
  const Space* s = sys->space;
  using C = typename C2<R>::T;
  const typename ST2<R>::T* st;
  const R* s0v;
  if constexpr (sizeof(R) == 4) { st = sys->st32.p; s0v = sys->s0_32.p; }
  else { st = sys->st64.p; s0v = sys->s0_64.p; }
  const bool has_s0 = (sys->which & SUCEM_MAT_S0) && s->nnz0 > 0;
  const int32_t* s0p = has_s0 ? s->s0_ptr.p : nullptr;
  R cs = (R)k.s, ct = (R)k.t, c0 = (R)k.s0;
  int n = (int)s->ndofs;
#define SYS_ARGS n, G, s->row_ptr.p, s->col.p, st, s0p, s->s0_col.p, s0v, cs, ct, c0, mask, \
                 (const C*)x, (C*)y, (const C*)w, rs
  if (epi == 0) SUCEM_LAUNCH((k_sys_spmv<R, L, U, 0>), G, TB, 0, c->stream, SYS_ARGS);
  else if (epi == 1) SUCEM_LAUNCH((k_sys_spmv<R, L, U, 1>), G, TB, 0, c->stream, SYS_ARGS);
  else SUCEM_LAUNCH((k_sys_spmv<R, L, U, 2>), G, TB, 0, c->stream, SYS_ARGS);
#undef SYS_ARGS
}

template <typename R, int L, int U>
void launch_csr(Context* c, const Csr* A, const uint8_t* mask, const void* x, void* y,
                const void* w, int epi, const RedSite& rs, int G) {
  using C = typename C2<R>::T;
  const C* val;
  if constexpr (sizeof(R) == 4) val = A->val32.p; else val = A->val64.p;
  int n = (int)A->n;
#define CSR_ARGS n, G, A->row_ptr.p, A->col.p, val, mask, (const C*)x, (C*)y, (const C*)w, rs
  if (epi == 0) SUCEM_LAUNCH((k_csr_spmv<R, L, U, 0>), G, TB, 0, c->stream, CSR_ARGS);
  else if (epi == 1) SUCEM_LAUNCH((k_csr_spmv<R, L, U, 1>), G, TB, 0, c->stream, CSR_ARGS);
  else SUCEM_LAUNCH((k_csr_spmv<R, L, U, 2>), G, TB, 0, c->stream, CSR_ARGS);
#undef CSR_ARGS
}

template <typename R, int L>
void sys_dispatch_u(Context* c, const System* sys, Coef k, const uint8_t* mask, const void* x,
                    void* y, const void* w, int epi, const RedSite& rs, int G, int U) {
  if (U <= 1) launch_sys<R, L, 1>(c, sys, k, mask, x, y, w, epi, rs, G);
  else if (U == 2) launch_sys<R, L, 2>(c, sys, k, mask, x, y, w, epi, rs, G);
  else launch_sys<R, L, 4>(c, sys, k, mask, x, y, w, epi, rs, G);
}

template <typename R, int L>
void csr_dispatch_u(Context* c, const Csr* A, const uint8_t* mask, const void* x, void* y,
                    const void* w, int epi, const RedSite& rs, int G, int U) {
  if (U <= 1) launch_csr<R, L, 1>(c, A, mask, x, y, w, epi, rs, G);
  else if (U == 2) launch_csr<R, L, 2>(c, A, mask, x, y, w, epi, rs, G);
  else launch_csr<R, L, 4>(c, A, mask, x, y, w, epi, rs, G);
}

}  // namespace

int lanes_for(int64_t n, int64_t nnz) {
  double avg = n ? (double)nnz / (double)n : 1.0;
  if (avg <= 6) return 4;
  if (avg <= 20) return 8;
  if (avg <= 48) return 16;
  return 32;
}

int op_grid(Context* c, int64_t n, int per_cu) {
  // a multiple of 8 (one share per XCD); at least a few rows per block
  int64_t want = (int64_t)c->num_cus * per_cu;
  int64_t g = (n + 63) / 64;
  if (g > want) g = want;
  g = (g + 7) / 8 * 8;
  if (g < 8) g = 8;
  return (int)g;
}

SpmvCfg default_cfg(Context* c, int64_t n, int64_t nnz) {
  SpmvCfg f;
  f.L = lanes_for(n, nnz);
  double avg = n ? (double)nnz / (double)n : 1.0;
  int u = (int)std::ceil(avg / f.L);
  f.U = u <= 1 ? 1 : (u == 2 ? 2 : 4);
  f.G = op_grid(c, n, 16);
  return f;
}

template <typename R>
void sys_apply(Context* c, const System* sys, Coef k, const uint8_t* mask, const void* x,
               void* y, const void* w, int epi, const RedSite& rs, const SpmvCfg& f) {
  if (f.mf == 2) {
    pt_apply<R>(c, sys, k, mask, x, y, w, epi, rs, f);
    return;
  }
  if (f.mf) {
    SUCEM_CHECK(sys->mf_ready, SUCEM_ERR_INVALID, "matrix-free operator not prepared");
    mf_apply<R>(c, sys, k, mask, x, y, w, epi, rs, f);
    return;
  }
  switch (f.L) {
    case 4: sys_dispatch_u<R, 4>(c, sys, k, mask, x, y, w, epi, rs, f.G, f.U); break;
    case 8: sys_dispatch_u<R, 8>(c, sys, k, mask, x, y, w, epi, rs, f.G, f.U); break;
    case 16: sys_dispatch_u<R, 16>(c, sys, k, mask, x, y, w, epi, rs, f.G, f.U); break;
    default: sys_dispatch_u<R, 32>(c, sys, k, mask, x, y, w, epi, rs, f.G, f.U); break;
  }
}
template void sys_apply<float>(Context*, const System*, Coef, const uint8_t*, const void*, void*,
                               const void*, int, const RedSite&, const SpmvCfg&);
template void sys_apply<double>(Context*, const System*, Coef, const uint8_t*, const void*,
                                void*, const void*, int, const RedSite&, const SpmvCfg&);

template <typename R>
void csr_apply(Context* c, const Csr* A, const uint8_t* mask, const void* x, void* y,
               const void* w, int epi, const RedSite& rs, const SpmvCfg& f) {
  switch (f.L) {
    case 4: csr_dispatch_u<R, 4>(c, A, mask, x, y, w, epi, rs, f.G, f.U); break;
    case 8: csr_dispatch_u<R, 8>(c, A, mask, x, y, w, epi, rs, f.G, f.U); break;
    case 16: csr_dispatch_u<R, 16>(c, A, mask, x, y, w, epi, rs, f.G, f.U); break;
    default: csr_dispatch_u<R, 32>(c, A, mask, x, y, w, epi, rs, f.G, f.U); break;
  }
}
template void csr_apply<float>(Context*, const Csr*, const uint8_t*, const void*, void*,
                               const void*, int, const RedSite&, const SpmvCfg&);
template void csr_apply<double>(Context*, const Csr*, const uint8_t*, const void*, void*,
                                const void*, int, const RedSite&, const SpmvCfg&);

template <typename R>
SpmvCfg sys_cfg(Context* c, const System* sys) {
  const SpmvCfg& t = sys->cfg[sizeof(R) == 4 ? 0 : 1];
  if (t.L > 0 || t.mf) return t;
  // Nothing tuned or set: the patch operator (what the autotuner picks on the driven
  // cubes, 8-10x faster than the assembled CSR operator), with the patch size it picks
  // (FP64 order 2: 192 cells, else 256), when the space suits it; the choice is kept on
  // the system (its maps and coefficients are lazily built caches of the system).
  System* ms = const_cast<System*>(sys);
  const int cb = (sys->space->order == 2 && sizeof(R) == 8) ? 192 : 256;
  if (sys->space->nloc == 6 || sys->space->nloc == 20) {
    if (pt_prepare(c, ms, cb) && pt_fits<R>(sys->space, cb)) {
      const int64_t npd = sys->space->pt[cb == 192 ? 1 : 2].npd;
      SpmvCfg f = pt_cfg(sys->space, cb, op_grid(c, std::max<int64_t>(npd, 1), 16));
      ms->cfg[sizeof(R) == 4 ? 0 : 1] = f;
      return f;
    }
  }
  return default_cfg(c, sys->space->ndofs, sys->space->nnz);
}
template SpmvCfg sys_cfg<float>(Context*, const System*);
template SpmvCfg sys_cfg<double>(Context*, const System*);

SpmvCfg csr_cfg(Context* c, const Csr* A) {
  return A->cfg.L > 0 ? A->cfg : default_cfg(c, A->n, A->nnz);
}

// Autotuner: time every candidate of the fused operator with HIP events on the
// context's stream and keep the fastest.  Inside a solve the operator's inputs have been
// pushed out of the caches by the other kernels of the iteration, so every timed
// application here follows a write of a buffer twice the size of the last-level cache
// (256 MB on MI355X); back-to-back applications would be served from that cache and
// favour a different form.
template <typename R>
SpmvCfg tune_sys(Context* c, System* sys, int reps) {
  const Space* s = sys->space;
  using C = typename C2<R>::T;
  DBuf<C> x(s->ndofs), y(s->ndofs);
  DBuf<char> flush((size_t)512 << 20);
  SUCEM_HIP(hipMemsetAsync(x.p, 0, x.bytes(), c->stream));
  hipEvent_t e0, e1;
  SUCEM_HIP(hipEventCreate(&e0));
  SUCEM_HIP(hipEventCreate(&e1));
  Coef k;
  k.s = 1.0; k.t = -1.0; k.s0 = 1.0;
  RedSite none{};
  SpmvCfg best = default_cfg(c, s->ndofs, s->nnz);
  best.ms = 1e30;
  auto time_cfg = [&](SpmvCfg& f) {
    sys_apply<R>(c, sys, k, nullptr, x.p, y.p, nullptr, 0, none, f);  // warm (first-use set-up)
    double total = 0;
    for (int i = 0; i < reps; ++i) {
      SUCEM_HIP(hipMemsetAsync(flush.p, i & 0xff, flush.bytes(), c->stream));
      SUCEM_HIP(hipEventRecord(e0, c->stream));
      sys_apply<R>(c, sys, k, nullptr, x.p, y.p, nullptr, 0, none, f);
      SUCEM_HIP(hipEventRecord(e1, c->stream));
      SUCEM_HIP(hipEventSynchronize(e1));
      float ms = 0;
      SUCEM_HIP(hipEventElapsedTime(&ms, e0, e1));
      total += ms;
    }
    f.ms = total / reps;
    if (f.ms < best.ms) best = f;
  };
  const int Ls[] = {4, 8, 16, 32};
  const int Us[] = {1, 2, 4};
  const int per_cu[] = {4, 8, 16, 32};
  const double avg = (double)s->nnz / (double)s->ndofs;
  for (int L : Ls)
    for (int U : Us) {
      if (L * U > 4 * avg + 8 || L * U * 4 < avg) continue;
      for (int pc : per_cu) {
        SpmvCfg f;
        f.L = L; f.U = U; f.G = op_grid(c, s->ndofs, pc);
        time_cfg(f);
      }
    }
  // matrix-free candidates: element products + row gather
  if (s->nloc == 6 || s->nloc == 20) {
    mf_prepare(c, sys);
    // element products whole (1), half the rows at a time (2), on the matrix cores (3:
    // order 2, FP32)
    for (int split : {1, 2, 3})
      for (int lanes : {0, 1, 4})
        for (int pc : per_cu) {
          if (split == 3 && !(sizeof(R) == 4 && s->order == 2)) continue;
          SpmvCfg f;
          f.mf = 1; f.L = lanes; f.U = split; f.G = op_grid(c, s->ndofs, pc);
          time_cfg(f);
        }
    // patch operator: element products summed per patch of cells in LDS + shared-DOF sums
    for (int cb : {128, 192, 256}) {
      if ((s->order == 1 && cb != 256) || !pt_prepare(c, sys, cb) || !pt_fits<R>(s, cb)) continue;
      const int64_t npd = s->pt[cb == 128 ? 0 : (cb == 192 ? 1 : 2)].npd;
      for (int pc : {4, 16}) {
        SpmvCfg f = pt_cfg(s, cb, op_grid(c, std::max<int64_t>(npd, 1), pc));
        try {
          time_cfg(f);
        } catch (const Error&) {
          break;   // a launch the device refuses (LDS per workgroup): not a candidate
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  sys->cfg[sizeof(R) == 4 ? 0 : 1] = best;
  return best;
}
template SpmvCfg tune_sys<float>(Context*, System*, int);
template SpmvCfg tune_sys<double>(Context*, System*, int);

template <typename R>
void sys_dinv(Context* c, const System* sys, Coef k, const uint8_t* mask, void* dinv) {
  const Space* s = sys->space;
  const bool has_s0 = (sys->which & SUCEM_MAT_S0) && s->nnz0 > 0;
  if constexpr (sizeof(R) == 4)
    SUCEM_LAUNCH((k_sys_dinv<float>), grid_for(s->ndofs, TB), TB, 0, c->stream, s->ndofs,
                 s->diag.p, sys->st32.p, has_s0 ? s->s0_diag.p : nullptr, sys->s0_32.p, k.s, k.t,
                 k.ti, k.s0, mask, (float2*)dinv);
  else
    SUCEM_LAUNCH((k_sys_dinv<double>), grid_for(s->ndofs, TB), TB, 0, c->stream, s->ndofs,
                 s->diag.p, sys->st64.p, has_s0 ? s->s0_diag.p : nullptr, sys->s0_64.p, k.s, k.t,
                 k.ti, k.s0, mask, (double2*)dinv);
}
template void sys_dinv<float>(Context*, const System*, Coef, const uint8_t*, void*);
template void sys_dinv<double>(Context*, const System*, Coef, const uint8_t*, void*);

template <typename R>
void csr_dinv(Context* c, const Csr* A, const uint8_t* mask, bool identity, void* dinv, int64_t n) {
  using C = typename C2<R>::T;
  if (identity) {
    SUCEM_LAUNCH((k_csr_identity_dinv<R>), grid_for(n, TB), TB, 0, c->stream, n, mask, (C*)dinv);
    return;
  }
  const C* val;
  if constexpr (sizeof(R) == 4) val = A->val32.p; else val = A->val64.p;
  SUCEM_LAUNCH((k_csr_dinv<R>), grid_for(A->n, TB), TB, 0, c->stream, A->n, A->diag.p, val, mask,
               (C*)dinv);
}
template void csr_dinv<float>(Context*, const Csr*, const uint8_t*, bool, void*, int64_t);
template void csr_dinv<double>(Context*, const Csr*, const uint8_t*, bool, void*, int64_t);

template <typename R>
void system_apply(Context* c, const System* sys, Coef k, bool eliminate_pec, int64_t nvec,
                  const void* x, void* y) {
  const Space* s = sys->space;
  SUCEM_CHECK(sys->st64.p, SUCEM_ERR_INVALID, "S and T must be assembled");
  const SpmvCfg f = sys_cfg<R>(c, sys);
  const size_t bytes = sizeof(typename C2<R>::T);
  RedSite none{};
  for (int64_t v = 0; v < nvec; ++v)
    sys_apply<R>(c, sys, k, eliminate_pec ? s->pec.p : nullptr,
                 (const char*)x + v * s->ndofs * bytes, (char*)y + v * s->ndofs * bytes, nullptr,
                 0, none, f);
}
template void system_apply<float>(Context*, const System*, Coef, bool, int64_t, const void*,
                                  void*);
template void system_apply<double>(Context*, const System*, Coef, bool, int64_t, const void*,
                                   void*);

void csr_apply64(Context* c, const Csr* A, const double2* x, double2* y) {
  RedSite none{};
  csr_apply<double>(c, A, nullptr, x, y, nullptr, 0, none, csr_cfg(c, A));
}

void locate_points(Context* c, const Space* s, int64_t n, const double* pts_d, int32_t* cell_d) {
  const Mesh* m = s->mesh;
  // initialise to "not found"
  std::vector<int32_t> init(n, 0x7fffffff);
  h2d(c, cell_d, init.data(), n);
  SUCEM_LAUNCH(k_locate, grid_for(m->nc, TB), TB, 0, c->stream, m->nc, m->xyz.p, m->cells.p, n,
               pts_d, 1e-12, cell_d);
  c->sync();
}

void rhs_point_sources(Context* c, const Space* s, int64_t nsrc, const double* pos_d,
                       const double2* val_d, const int32_t* rhs_d, const int32_t* cell_d,
                       int64_t nrhs, double2 scale, double2* b_d) {
  const Mesh* m = s->mesh;
  unsigned g = (unsigned)(nrhs < 65535 ? nrhs : 65535);
  if (g == 0 || nsrc == 0) return;
  SUCEM_LAUNCH(k_rhs, g, 64, 0, c->stream, nsrc, pos_d, val_d, rhs_d, cell_d, nrhs, scale,
               m->xyz.p, m->cells.p, s->cell_dofs.p, s->ref_basis.p, s->nloc, s->ndofs, b_d);
}

void basis_at_points(Context* c, const Space* s, int64_t n, const double* pts_d,
                     const int32_t* cell_d, double* vals_d) {
  const Mesh* m = s->mesh;
  if (n == 0) return;
  SUCEM_LAUNCH(k_basis_at, grid_for(n, 64), 64, 0, c->stream, n, pts_d, cell_d, m->xyz.p,
               m->cells.p, s->ref_basis.p, s->nloc, vals_d);
}

// ---- allocation-free point-source right-hand sides -----------------------------------
// Staging of sucem_rhs_point_sources_noalloc lives in a fixed array of device memory (one
// copy per GPU, held until the process ends) instead of per-call allocations.  In host
// code the array's name only stands for the symbol: its device address on each GPU comes
// from hipGetSymbolAddress (noalloc_init, called when a context is created on that GPU),
// and only that address is handed to kernels and copies.  Every call on a GPU shares the
// array: calls that overlap on one GPU race.
namespace {

constexpr int64_t kNA = kNoAllocSources;
constexpr size_t kNaPos = 0;                                   // double[3 kNA]
constexpr size_t kNaVal = kNaPos + 24 * kNA;                   // double2[3 kNA]
constexpr size_t kNaRhs64 = kNaVal + 48 * kNA;                 // int64[kNA]
constexpr size_t kNaRhs = kNaRhs64 + 8 * kNA;                  // int32[kNA]
constexpr size_t kNaCell = kNaRhs + 4 * kNA;                   // int32[kNA]
constexpr size_t kNaFlag = kNaCell + 4 * kNA;                  // int32: a point is outside
constexpr size_t kNaBytes = kNaFlag + 256;
static_assert(kNaVal % 256 == 0 && kNaRhs64 % 256 == 0 && kNaRhs % 256 == 0 &&
                  kNaCell % 256 == 0 && kNaFlag % 256 == 0,
              "staging sections are 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

std::mutex g_na_mu;
constexpr int kNaMaxDevices = 64;
unsigned char* g_na_ptr[kNaMaxDevices] = {};

// rhs32 = rhs64 (0 without indices), cell = "not found", the outside flag cleared
__global__ void k_na_prep(int64_t n, const int64_t* __restrict__ rhs64, int32_t* __restrict__ rhs,
                          int32_t* __restrict__ cell, int32_t* __restrict__ outside) {
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t0 == 0) *outside = 0;
  for (int64_t i = t0; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    rhs[i] = rhs64 ? (int32_t)rhs64[i] : 0;
    cell[i] = 0x7fffffff;
  }
}

__global__ void k_na_outside(int64_t n, const int32_t* __restrict__ cell,
                             int32_t* __restrict__ outside) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (cell[i] == 0x7fffffff) atomicOr(outside, 1);
}

}  // namespace

void noalloc_init(int device) {
  std::lock_guard<std::mutex> hold(g_na_mu);
  SUCEM_CHECK(device >= 0 && device < kNaMaxDevices, SUCEM_ERR_UNSUPPORTED,
              "device index beyond the allocation-free staging table");
  if (g_na_ptr[device]) return;
  void* p = nullptr;
  SUCEM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  SUCEM_CHECK(p != nullptr, SUCEM_ERR_HIP, "no device address for the staging array");
  g_na_ptr[device] = (unsigned char*)p;
}

void rhs_point_sources_noalloc(Context* c, const Space* s, int64_t nsrc, const double* pos_h,
                               const double* val_h, const int64_t* rhs_h, int64_t nrhs,
                               double2 scale, double2* b_d) {
  SUCEM_CHECK(nsrc >= 0 && nsrc <= kNA, SUCEM_ERR_INVALID,
              "the allocation-free point-source call takes at most 4096 sources");
  unsigned char* base;
  {
    std::lock_guard<std::mutex> hold(g_na_mu);
    base = (c->device >= 0 && c->device < kNaMaxDevices) ? g_na_ptr[c->device] : nullptr;
  }
  SUCEM_CHECK(base != nullptr, SUCEM_ERR_INTERNAL,
              "allocation-free staging not initialised on this device");
  if (nsrc == 0 || nrhs == 0) return;
  double* pos = (double*)(base + kNaPos);
  double2* val = (double2*)(base + kNaVal);
  int64_t* rhs64 = (int64_t*)(base + kNaRhs64);
  int32_t* rhs = (int32_t*)(base + kNaRhs);
  int32_t* cell = (int32_t*)(base + kNaCell);
  int32_t* outside = (int32_t*)(base + kNaFlag);
  SUCEM_HIP(hipMemcpyAsync(pos, pos_h, sizeof(double) * 3 * nsrc, hipMemcpyHostToDevice,
                           c->stream));
  SUCEM_HIP(hipMemcpyAsync(val, val_h, sizeof(double2) * 3 * nsrc, hipMemcpyHostToDevice,
                           c->stream));
  if (rhs_h)
    SUCEM_HIP(hipMemcpyAsync(rhs64, rhs_h, sizeof(int64_t) * nsrc, hipMemcpyHostToDevice,
                             c->stream));
  SUCEM_LAUNCH(k_na_prep, grid_for(nsrc, 64), 64, 0, c->stream, nsrc,
               rhs_h ? (const int64_t*)rhs64 : nullptr, rhs, cell, outside);
  const Mesh* m = s->mesh;
  SUCEM_LAUNCH(k_locate, grid_for(m->nc, TB), TB, 0, c->stream, m->nc, m->xyz.p, m->cells.p,
               nsrc, (const double*)pos, 1e-12, cell);
  SUCEM_LAUNCH(k_na_outside, grid_for(nsrc, 64), 64, 0, c->stream, nsrc, (const int32_t*)cell,
               outside);
  int32_t out_h = 0;
  SUCEM_HIP(hipMemcpyAsync(&out_h, outside, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  c->sync();
  SUCEM_CHECK(out_h == 0, SUCEM_ERR_OUTSIDE, "point source is outside the mesh");
  rhs_point_sources(c, s, nsrc, pos, val, rhs, cell, nrhs, scale, b_d);
  c->sync();
}

}  // namespace sucem
