// Resonant-cavity eigen solver (replaces DefaultEigenSolver.solve_problem,
// ProblemConfigurations/EMVectorWaveEigenproblem.py:61-105: LU of S - sigma M and
// ARPACK shift-invert).
//
// Thick-restart Lanczos on OP = (S - sigma T)^-1 T, self-adjoint in the T inner
// product, on the PEC-eliminated pencil (rows/columns of the DOFs on boundary
// facets removed).  The wanted modes are the largest positive Ritz values
// nu = 1/(lambda - sigma) > 0, i.e. the lowest k^2 above sigma; the gradient null
// space of the curl-curl operator maps to nu = -1/sigma < 0 and is never wanted.
// Inner solves use the device COCG with the auxiliary-space preconditioner: FP64, or FP32
// refinement cycles (FP32 recursions, each restarted from the residual recomputed with the
// FP64 operator, x accumulated in FP64, a cycle cut short when its recursion stagnates)
// that reach the same inner tolerance; orthogonalisation, the projected matrix and its
// eigen-decomposition are FP64, and each returned k^2 is the FP64 Rayleigh quotient
// x^T S x / x^T T x of its Ritz vector, with the FP64 residual ||S x - k^2 T x||.
//
// Block size 2: the basis grows by two vectors per step, and the two inner solves of a
// step run as one paired solve (the two real right-hand sides packed as the real and
// imaginary parts of one complex vector, two independent real COCG recursions over the
// same operator and preconditioner passes).  The complex kernels then do no work on
// zero imaginary parts.  Each new vector is re-orthogonalised against the whole basis
// (its partner included), so the projected matrix is banded with two sub-diagonals and
// is formed from the Gram-Schmidt coefficients as for block size 1; the thick restart
// keeps Ritz vectors plus both residual vectors.  On a dielectric cavity (16 x 8 x 12
// boxes, order 2, 20 modes) block 2 took 57 steps (114 vectors) against 96 for block 1.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "internal.hpp"
#include "kernels.hpp"

namespace sucem {
namespace {

constexpr int TB = 256;

struct Range { int start, end; };
__device__ __forceinline__ Range rows_of(int n, int G) {
  int lb = xcd_block(G);
  int chunk = (n + G - 1) / G;
  Range r;
  r.start = lb * chunk;
  r.end = min(n, r.start + chunk);
  return r;
}

// deterministic pseudo-random start vector (splitmix64), zero on masked DOFs
__global__ void k_start_vector(int64_t n, unsigned long long seed, const uint8_t* mask,
                               double2* v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    double u = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    v[i] = make_double2((mask && mask[i]) ? 0.0 : u, 0.0);
  }
}

// part[i*G + lb] = sum_r Re(conj(a_i[r]) y[r]) for i < m  (vectors a_i = base + i*n).
// The basis vectors are taken MB at a time, so y is read once per MB of them.
constexpr int MB = 8;
__global__ __launch_bounds__(TB) void k_multidot(int n, int G, int m, const double2* base,
                                                 const double2* y, double2* part) {
  Range rr = rows_of(n, G);
  for (int i0 = 0; i0 < m; i0 += MB) {
    const int cnt = min(MB, m - i0);
    double acc[MB];
#pragma unroll
    for (int k = 0; k < MB; ++k) acc[k] = 0.0;
    for (int r = rr.start + threadIdx.x; r < rr.end; r += TB) {
      const double2 yr = y[r];
#pragma unroll
      for (int k = 0; k < MB; ++k) {
        if (k < cnt) {
          const double2 a = base[(int64_t)(i0 + k) * n + r];
          acc[k] += a.x * yr.x + a.y * yr.y;
        }
      }
    }
    for (int k = 0; k < cnt; ++k) {
      const double2 t = block_sum<TB>(make_double2(acc[k], 0.0));
      if (threadIdx.x == 0) part[(int64_t)(i0 + k) * G + xcd_block(G)] = t;
    }
  }
}

// y -= sum_i h[i] v_i
__global__ void k_multiaxpy(int n, int G, int m, const double* h, const double2* base,
                            double2* y) {
  Range rr = rows_of(n, G);
  for (int r = rr.start + threadIdx.x; r < rr.end; r += TB) {
    double2 acc = y[r];
    for (int i = 0; i < m; ++i) {
      const double2 v = base[(int64_t)i * n + r];
      acc.x -= h[i] * v.x;
      acc.y -= h[i] * v.y;
    }
    y[r] = acc;
  }
}

// out_c = sum_i in_i * S[i*ld + c] for c < k (restart combination)
__global__ void k_combine(int n, int G, int m, int k, const double* S, int ld, const double2* in,
                          double2* out) {
  Range rr = rows_of(n, G);
  for (int r = rr.start + threadIdx.x; r < rr.end; r += TB) {
    for (int c = 0; c < k; ++c) {
      double2 acc = make_double2(0, 0);
      for (int i = 0; i < m; ++i) {
        const double2 v = in[(int64_t)i * n + r];
        const double s = S[i * ld + c];
        acc.x += s * v.x;
        acc.y += s * v.y;
      }
      out[(int64_t)c * n + r] = acc;
    }
  }
}

__global__ void k_scale_copy(int n, int G, double s, const double2* in, double2* out) {
  Range rr = rows_of(n, G);
  for (int r = rr.start + threadIdx.x; r < rr.end; r += TB)
    out[r] = make_double2(in[r].x * s, in[r].y * s);
}

// b = a0 + i a1 for two real vectors a0, a1 (held as complex with zero imaginary parts)
__global__ void k_pack2(int n, int G, const double2* a0, const double2* a1, double2* b) {
  Range rr = rows_of(n, G);
  for (int r = rr.start + threadIdx.x; r < rr.end; r += TB) b[r] = make_double2(a0[r].x, a1[r].x);
}

// y0 = Re y, y1 = Im y (as complex vectors with zero imaginary parts)
__global__ void k_unpack2(int n, int G, const double2* y, double2* y0, double2* y1) {
  Range rr = rows_of(n, G);
  for (int r = rr.start + threadIdx.x; r < rr.end; r += TB) {
    const double2 v = y[r];
    y0[r] = make_double2(v.x, 0.0);
    y1[r] = make_double2(v.y, 0.0);
  }
}

// Symmetric eigen-decomposition of a small dense matrix (cyclic Jacobi), ascending.
void jacobi_eig(std::vector<double> A, int n, std::vector<double>& w, std::vector<double>& V) {
  V.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        tot += A[i * n + j] * A[i * n + j];
        if (i != j) off += A[i * n + j] * A[i * n + j];
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        double app = A[p * n + p], aqq = A[q * n + q];
        double theta = (aqq - app) / (2 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return A[a * n + a] < A[b * n + b]; });
  w.resize(n);
  std::vector<double> Vs(n * n);
  for (int c = 0; c < n; ++c) {
    w[c] = A[idx[c] * n + idx[c]];
    for (int r = 0; r < n; ++r) Vs[r * n + c] = V[r * n + idx[c]];
  }
  V.swap(Vs);
}

}  // namespace

void eigs_system(Context* c, const System* sys, double sigma, int nev, const EigOpts& o,
                 double* lam_h, double* resid_h, double2* vecs_d, EigStats& st) {
  const Space* s = sys->space;
  SUCEM_CHECK(sys->st64.p, SUCEM_ERR_INVALID, "S and T must be assembled");
  const int n = (int)s->ndofs;
  const int64_t nfree = s->ndofs - s->npec;
  SUCEM_CHECK(nev > 0 && nev < nfree, SUCEM_ERR_INVALID, "nev must be in [1, free DOFs)");
  // block size: 2 where the inner solves can run paired and the basis fits, else 1
  int bs = o.block == 1 ? 1 : 2;
  if (o.block == 0 && !paired_solves_available(c, sys, o.inner_precision)) bs = 1;
  SUCEM_CHECK(bs == 1 || paired_solves_available(c, sys, o.inner_precision),
              SUCEM_ERR_UNSUPPORTED,
              "Lanczos block size 2 needs paired inner solves (the patch operator)");
  // default basis: max(2 nev + 1, 20) for block 1; for block 2 also at least nev + 20, so
  // that a restart cycle adds 10 or more vectors (CPU, dielectric cavity 16 x 8 x 12, order
  // 2, exact inner solves, block 2: 8 modes 52 / 39 / 39 / 36 steps at a basis of 20 / 28 /
  // 30 / 40, 10 modes 53 / 40 at 22 / 30, 4 modes 26 / 27 at 20 / 24)
  auto basis_size = [&](int b) {
    int mb = o.ncv > 0 ? o.ncv : std::max(2 * nev + b, b == 2 ? nev + 20 : 20);
    // block 2 holds m + 2 T-orthonormal vectors (block 1's last vector ends the cycle as an
    // invariant subspace when m = nfree)
    mb = (int)std::min<int64_t>(mb, b == 1 ? nfree : nfree - b);
    return mb - mb % b;
  };
  int m = basis_size(bs);
  if (bs == 2 && o.block == 0 && m < nev + 2) {
    bs = 1;
    m = basis_size(1);
  }
  SUCEM_CHECK(m >= nev + bs, SUCEM_ERR_INVALID, "ncv must exceed nev");
  auto t0 = std::chrono::steady_clock::now();
  st = EigStats();
  st.block = bs;
  // progress lines on stderr: SUCEM_LOG=1 per restart cycle, 2 also per Lanczos step
  const char* lg = std::getenv("SUCEM_LOG");
  const int log_level = lg ? std::atoi(lg) : 0;
  auto elapsed_s = [&] {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  const int G = [&] {
    int64_t g = std::min<int64_t>((int64_t)c->num_cus * 4, (n + 255) / 256);
    g = (g + 7) / 8 * 8;
    return (int)std::max<int64_t>(g, 8);
  }();
  const uint8_t* mask = s->pec.p;

  // basis, T times basis, restart scratch and work vectors in the context's workspace
  // (slots 32..43): repeated eigen solves on one context allocate nothing after the first
  const int mb = m + bs;   // basis vectors including the residual block
  struct Buf { double2* p; };
  const Buf V{ws_take<double2>(c, 32, (int64_t)mb * n)};
  const Buf TV{ws_take<double2>(c, 33, (int64_t)mb * n)};
  const Buf W{ws_take<double2>(c, 34, (int64_t)(m + 1) * n)};
  const Buf y{ws_take<double2>(c, 35, n)}, ty{ws_take<double2>(c, 36, n)};
  const Buf part{ws_take<double2>(c, 37, (int64_t)mb * G)};
  struct { double* p; } hdev{ws_take<double>(c, 38, mb)};
  // small combination matrices of the restarts
  double* Sd = ws_take<double>(c, 39, (int64_t)(m + 1) * (m + 1));
  // block 2: the second solution and the packed right-hand side and solution
  const Buf y1{ws_take<double2>(c, 41, bs == 2 ? n : 0)};
  const Buf bp{ws_take<double2>(c, 42, bs == 2 ? n : 0)};
  const Buf yp{ws_take<double2>(c, 43, bs == 2 ? n : 0)};
  std::vector<double> H((size_t)mb * mb, 0.0);
  auto h = [&](int i, int j) -> double& { return H[(size_t)i * mb + j]; };

  Coef Tk;  // T
  Tk.s = 0; Tk.t = 1; Tk.s0 = 0;
  Coef Sk;  // S
  Sk.s = 1; Sk.t = 0; Sk.s0 = 0;
  Coef Ak;  // S - sigma T
  Ak.s = 1; Ak.t = -sigma; Ak.s0 = 0;

  auto vec = [&](const Buf& B, int i) { return B.p + (int64_t)i * n; };
  auto multidot = [&](const double2* base, int cnt, const double2* v, std::vector<double>& out) {
    SUCEM_LAUNCH(k_multidot, G, TB, 0, c->stream, n, G, cnt, base, v, part.p);
    std::vector<double2> hp((int64_t)cnt * G);
    SUCEM_HIP(hipMemcpyAsync(hp.data(), part.p, sizeof(double2) * cnt * G, hipMemcpyDeviceToHost,
                             c->stream));
    c->sync();
    out.assign(cnt, 0.0);
    for (int i = 0; i < cnt; ++i)
      for (int b = 0; b < G; ++b) out[i] += hp[(int64_t)i * G + b].x;
  };
  auto apply = [&](Coef k, const double2* x, double2* yv) {
    system_apply<double>(c, sys, k, true, 1, x, yv);
  };
  // wv -= its T-projection onto the basis vectors 0..top-1 (classical Gram-Schmidt, twice;
  // with col >= 0 the coefficients are added to column col of H), then
  // V[top] = wv / |wv|_T and TV[top] = T V[top]; returns |wv|_T (nothing is stored when
  // it is not finite and positive)
  auto orth_normalize = [&](double2* wv, int top, int col) {
    for (int pass = 0; pass < 2 && top > 0; ++pass) {
      std::vector<double> hh;
      multidot(TV.p, top, wv, hh);
      if (col >= 0)
        for (int i = 0; i < top; ++i) h(i, col) += hh[i];
      SUCEM_HIP(hipMemcpyAsync(hdev.p, hh.data(), sizeof(double) * top, hipMemcpyHostToDevice,
                               c->stream));
      SUCEM_LAUNCH(k_multiaxpy, G, TB, 0, c->stream, n, G, top, hdev.p, V.p, wv);
      c->sync();
    }
    apply(Tk, wv, ty.p);
    std::vector<double> nn;
    multidot(wv, 1, ty.p, nn);
    const double b = std::sqrt(std::max(nn[0], 0.0));
    if (std::isfinite(b) && b > 0.0) {
      SUCEM_LAUNCH(k_scale_copy, G, TB, 0, c->stream, n, G, 1.0 / b, wv, vec(V, top));
      SUCEM_LAUNCH(k_scale_copy, G, TB, 0, c->stream, n, G, 1.0 / b, ty.p, vec(TV, top));
    }
    return b;
  };
  // V[top]: a fresh random vector T-orthonormal to the basis 0..top-1
  auto random_vector = [&](int top, unsigned long long salt) {
    SUCEM_LAUNCH(k_start_vector, grid_for(n, TB), TB, 0, c->stream, (int64_t)n,
                 (unsigned long long)o.seed + salt, mask, y.p);
    double n0 = 1.0;
    if (top > 0) {   // its T-norm before the projection, to judge what is left of it
      apply(Tk, y.p, ty.p);
      std::vector<double> nn;
      multidot(y.p, 1, ty.p, nn);
      n0 = std::sqrt(std::max(nn[0], 0.0));
    }
    const double b = orth_normalize(y.p, top, -1);
    SUCEM_CHECK(std::isfinite(b) && b > (top > 0 ? 1e-8 * n0 : 0.0), SUCEM_ERR_INVALID,
                "no start vector independent of the basis (too few free DOFs)");
  };
  SolveOpts so;
  so.method = SUCEM_METHOD_COCG;
  so.precision = o.inner_precision;
  so.precond = SUCEM_PRECOND_AUXILIARY;
  so.max_iter = o.inner_max_iter;
  so.check_every = 8;
  so.tol = o.inner_tol;
  so.eliminate_pec = 1;
  so.precond_shift = 0.0;   // a shifted preconditioner slows the cavity operator S - sigma T
  so.paired = bs == 2 ? 1 : 0;
  int fp32_failed_run = 0;

  // start block: random, T-orthonormal
  for (int a = 0; a < bs; ++a) random_vector(a, (unsigned long long)a * 0xD1B54A32D192ED03ull);
  int k = 0;            // kept vectors at the start of this cycle
  std::vector<double> theta, Sv;
  std::vector<int> wanted;
  bool done = false;
  for (int cycle = 0; cycle <= o.max_restarts && !done; ++cycle) {
    int mm = m;
    for (int j = k; j < m; j += bs) {
      SolveStats ist;
      if (bs == 2) {
        SUCEM_LAUNCH(k_pack2, G, TB, 0, c->stream, n, G, vec(TV, j), vec(TV, j + 1), bp.p);
        solve_system(c, sys, Ak, bp.p, yp.p, so, ist, nullptr, 0);
        SUCEM_LAUNCH(k_unpack2, G, TB, 0, c->stream, n, G, yp.p, y.p, y1.p);
      } else {
        solve_system(c, sys, Ak, vec(TV, j), y.p, so, ist, nullptr, 0);
      }
      st.inner_iterations += ist.iterations;
      st.lanczos_steps += bs;
      if (ist.stagnated) st.inner_stagnated += 1;
      else if (ist.status != 0) st.inner_failures += 1;
      // FP32 inner cycles that fail (and are finished in FP64) on two solves in a row will
      // fail on the rest: the remaining inner solves run in FP64 from the start
      if (so.precision == 32) {
        fp32_failed_run = ist.fp32_failed ? fp32_failed_run + 1 : 0;
        if (fp32_failed_run >= 2) {
          so.precision = 64;
          st.inner_fp64_from = st.lanczos_steps + 1;
          if (log_level >= 1)
            std::fprintf(stderr, "[sucem eigs] FP32 inner cycles failed on two solves in a "
                                 "row: FP64 inner solves from vector %d\n", st.inner_fp64_from);
        }
      }
      if (log_level >= 2)
        std::fprintf(stderr, "[sucem eigs] vectors %d: inner %d iterations, rel %.2e%s, %.1f s\n",
                     st.lanczos_steps, ist.iterations, ist.rel_residual,
                     ist.stagnated ? " (floor)" : "", elapsed_s());
      bool invariant = false;
      for (int a = 0; a < bs; ++a) {
        const int col = j + a, top = j + bs + a;
        double2* wv = a == 0 ? y.p : y1.p;
        const double b = orth_normalize(wv, top, col);
        h(top, col) = b;
        if (!(b > 1e-14 * std::fabs(h(col, col))) || !std::isfinite(b)) {
          if (bs == 1) {
            invariant = true;   // invariant subspace: stop the cycle here
            break;
          }
          // block 2: the direction is spent; a random vector orthogonal to the basis
          // takes its place, with a zero coefficient
          h(top, col) = 0.0;
          random_vector(top, 0x8CB92BA72F3D8DD7ull * (unsigned long long)(top + 1));
        }
      }
      if (invariant) {
        mm = j + 1;
        break;
      }
    }
    // Rayleigh-Ritz on the projected (symmetrised) matrix
    std::vector<double> P(mm * mm);
    for (int i = 0; i < mm; ++i)
      for (int j = 0; j < mm; ++j) {
        double a = i <= j ? h(i, j) : h(j, i);
        P[i * mm + j] = a;
      }
    jacobi_eig(P, mm, theta, Sv);
    // wanted: the nev largest positive theta
    wanted.clear();
    for (int i = mm - 1; i >= 0 && (int)wanted.size() < nev; --i)
      if (theta[i] > 0) wanted.push_back(i);
    SUCEM_CHECK((int)wanted.size() == nev, SUCEM_ERR_INVALID,
                "fewer positive Ritz values than requested modes (sigma above the spectrum?)");
    // residual couplings of Ritz vector w to the residual block: OP u_w - theta_w u_w =
    // sum_a R[a][w] V[mm + a]
    auto resid_coupling = [&](int a, int w) {
      double r = 0;
      for (int cc = 0; cc < mm; ++cc) r += h(mm + a, cc) * Sv[cc * mm + w];
      return r;
    };
    double worst = 0;
    bool conv = true;
    for (int w : wanted) {
      double r2 = 0;
      for (int a = 0; a < bs; ++a) {
        const double r = resid_coupling(a, w);
        r2 += r * r;
      }
      double rel = std::sqrt(r2) / std::fabs(theta[w]);
      worst = std::max(worst, rel);
      if (rel > o.tol) conv = false;
    }
    st.max_ritz_residual = worst;
    st.restarts = cycle;
    if (log_level >= 1)
      std::fprintf(stderr, "[sucem eigs] cycle %d: %d Lanczos vectors, %lld inner iterations, "
                   "worst Ritz residual %.2e, %.1f s\n", cycle, st.lanczos_steps,
                   (long long)st.inner_iterations, worst, elapsed_s());
    if (conv || mm < m || cycle == o.max_restarts) {
      done = true;
      st.status = conv ? 0 : 1;
      // Ritz vectors of the wanted values into W[0..nev)
      std::vector<double> Sw(mm * nev);
      for (int i = 0; i < mm; ++i)
        for (int q = 0; q < nev; ++q) Sw[i * nev + q] = Sv[i * mm + wanted[q]];
      h2d(c, Sd, Sw.data(), Sw.size());
      SUCEM_LAUNCH(k_combine, G, TB, 0, c->stream, n, G, mm, nev, Sd, nev, V.p, W.p);
      c->sync();
      break;
    }
    // thick restart: keep kk Ritz vectors with the largest theta plus the residual block
    // (kk leaves a multiple of the block size to the next cycle)
    int kk = std::min(m - bs, nev + (m - nev) / 2);
    kk -= (m - kk) % bs;
    std::vector<int> keep;
    for (int i = mm - 1; i >= 0 && (int)keep.size() < kk; --i) keep.push_back(i);
    std::vector<double> Sk_(mm * kk);
    for (int i = 0; i < mm; ++i)
      for (int q = 0; q < kk; ++q) Sk_[i * kk + q] = Sv[i * mm + keep[q]];
    std::vector<double> coupling((size_t)bs * kk);
    for (int a = 0; a < bs; ++a)
      for (int q = 0; q < kk; ++q) coupling[(size_t)a * kk + q] = resid_coupling(a, keep[q]);
    h2d(c, Sd, Sk_.data(), Sk_.size());
    SUCEM_LAUNCH(k_combine, G, TB, 0, c->stream, n, G, mm, kk, Sd, kk, V.p, W.p);
    SUCEM_HIP(hipMemcpyAsync(vec(V, 0), W.p, sizeof(double2) * (int64_t)kk * n,
                             hipMemcpyDeviceToDevice, c->stream));
    SUCEM_LAUNCH(k_combine, G, TB, 0, c->stream, n, G, mm, kk, Sd, kk, TV.p, W.p);
    SUCEM_HIP(hipMemcpyAsync(vec(TV, 0), W.p, sizeof(double2) * (int64_t)kk * n,
                             hipMemcpyDeviceToDevice, c->stream));
    SUCEM_HIP(hipMemcpyAsync(vec(V, kk), vec(V, mm), sizeof(double2) * (int64_t)bs * n,
                             hipMemcpyDeviceToDevice, c->stream));
    SUCEM_HIP(hipMemcpyAsync(vec(TV, kk), vec(TV, mm), sizeof(double2) * (int64_t)bs * n,
                             hipMemcpyDeviceToDevice, c->stream));
    c->sync();
    std::fill(H.begin(), H.end(), 0.0);
    // the kept Ritz vectors: diagonal theta and their coupling rows to the residual block;
    // the coupling columns are rebuilt by the re-orthogonalisation of the next steps
    for (int q = 0; q < kk; ++q) {
      h(q, q) = theta[keep[q]];
      for (int a = 0; a < bs; ++a) h(kk + a, q) = coupling[(size_t)a * kk + q];
    }
    k = kk;
  }
  // FP64 Rayleigh quotients and true residuals of the Ritz vectors
  std::vector<double> tscale(nev, 1.0);
  for (int q = 0; q < nev; ++q) {
    double2* x = vec(W, q);
    apply(Sk, x, y.p);
    apply(Tk, x, ty.p);
    std::vector<double> a, b;
    multidot(x, 1, y.p, a);
    multidot(x, 1, ty.p, b);
    double lam = a[0] / b[0];
    // r = S x - lam T x, relative to |lam| |T x|
    std::vector<double> hl(1, lam);
    SUCEM_HIP(hipMemcpyAsync(hdev.p, hl.data(), sizeof(double), hipMemcpyHostToDevice, c->stream));
    SUCEM_LAUNCH(k_multiaxpy, G, TB, 0, c->stream, n, G, 1, hdev.p, ty.p, y.p);
    std::vector<double> rr, tt;
    multidot(y.p, 1, y.p, rr);
    multidot(ty.p, 1, ty.p, tt);
    lam_h[q] = lam;
    if (resid_h) resid_h[q] = std::sqrt(rr[0]) / (std::fabs(lam) * std::sqrt(tt[0]));
    tscale[q] = 1.0 / std::sqrt(b[0]);
  }
  // ascending k^2
  std::vector<int> ord(nev);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return lam_h[a] < lam_h[b]; });
  std::vector<double> l2(nev), r2(nev);
  for (int q = 0; q < nev; ++q) {
    l2[q] = lam_h[ord[q]];
    r2[q] = resid_h ? resid_h[ord[q]] : 0;
  }
  for (int q = 0; q < nev; ++q) {
    lam_h[q] = l2[q];
    if (resid_h) resid_h[q] = r2[q];
  }
  // T-normalised modes in ascending k^2 (the Ritz vectors are still in W)
  if (vecs_d)
    for (int q = 0; q < nev; ++q)
      SUCEM_LAUNCH(k_scale_copy, G, TB, 0, c->stream, n, G, tscale[ord[q]], vec(W, ord[q]),
                   vecs_d + (int64_t)q * n);
  c->sync();
  st.nconv = st.status == 0 ? nev : 0;
  st.elapsed_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace sucem
