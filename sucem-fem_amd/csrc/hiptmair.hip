// Gradient-space (Hiptmair-type) additive preconditioner for the curl-curl systems.
//
//   M^-1 r = D^-1 r + G W G^T r,   W = diag(G^T A G)^-1,
//
// with D = diag(A) and G the discrete gradient from the Lagrange space of the
// same order (P1 vertices, or P2 vertices + edge midpoints) into the Nedelec
// space: G[i][j] = nodal dual i applied to grad(phi_j).  Jacobi alone leaves the
// gradient fields (the near-kernel of curl-curl, shifted to -k0^2 T by the mass
// term) almost untouched; the second term acts on exactly that subspace.  M is
// symmetric, so COCG stays applicable.  A = cs S + ct T + i c0 S0 and
// G^T S G = 0, so its diagonal is formed per solve from three stored diagonals.
//
// The auxiliary-space form adds the vector nodal term of the Hiptmair-Xu splitting:
//
//   M^-1 r = D^-1 r + G W G^T r + Pi V Pi^T r,   V = theta_v diag(Pi^T A Pi)^-1
//   (theta_v: see nodal_weight),
//
// Pi the nodal interpolation of continuous piecewise-linear vector fields into the
// Nedelec space: Pi[i][(v, d)] = lambda_v(x_i) (t_i)_d for the DOF i with dual point x_i
// and tangent t_i.  Jacobi and the gradient term leave the smooth non-gradient fields
// (the large-scale part of curl-curl) to the Krylov method; Pi^T A Pi acts on them as a
// vector Laplacian-like operator whose diagonal is a usable smoother.
#include <cmath>
#include <vector>

#include "internal.hpp"
#include "kernels.hpp"
#include "ops.hpp"
#include "refdata.hpp"

namespace sucem {
namespace {

constexpr int TB = 256;

// Contributions (ned dof, lag dof) of every cell's nonzero local gradient entries.
__global__ void k_grad_keys(int64_t nc, int nloc, int nlagl, int npairs, const int32_t* pairs,
                            const int32_t* cell_dofs, const int32_t* cells,
                            const int32_t* cell_edges, int64_t nv, int B, int transpose,
                            uint64_t* keys, uint32_t* vals) {
  const int64_t n = nc * npairs;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / npairs;
    const int p = (int)(t % npairs);
    const int i = pairs[2 * p], j = pairs[2 * p + 1];
    const uint64_t ned = (uint64_t)cell_dofs[c * nloc + i];
    const uint64_t lag = j < 4 ? (uint64_t)cells[4 * c + j]
                               : (uint64_t)(nv + cell_edges[6 * c + (j - 4)]);
    keys[t] = transpose ? ((lag << B) | ned) : ((ned << B) | lag);
    vals[t] = (uint32_t)t;
  }
}

// value of each unique entry = the local gradient entry of its first contribution; also
// packed with its column as (col << 4) | (k & 15), value = k / den
__global__ void k_grad_vals(int64_t nnz, const int32_t* seg, const uint32_t* perm, int npairs,
                            const double* pair_vals, const int32_t* pair_num,
                            const int32_t* colv, double* v64, float* v32, uint32_t* pk) {
  for (int64_t z = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; z < nnz;
       z += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t id = perm[seg[z]] % (uint32_t)npairs;
    const double v = pair_vals[id];
    v64[z] = v;
    v32[z] = (float)v;
    pk[z] = ((uint32_t)colv[z] << 4) | ((uint32_t)pair_num[id] & 15u);
  }
}

__device__ __forceinline__ int pk_col(uint32_t p) { return (int)(p >> 4); }
__device__ __forceinline__ int pk_num(uint32_t p) { return ((int)(p << 28)) >> 28; }  // signed

// One wavefront per Lagrange node j: out[3j..3j+2] = g_j^T {S, T, S0} g_j over the
// (unmasked) support of column j of G.
__global__ __launch_bounds__(TB) void k_grad_diag(int64_t nlag, const int32_t* gt_ptr,
                                                  const int32_t* gt_col, const double* gt_val,
                                                  const int32_t* rp, const int32_t* col,
                                                  const double2* st, const int32_t* s0p,
                                                  const int32_t* s0c, const double* s0v,
                                                  const uint8_t* mask, double* out) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t j = wave; j < nlag; j += nwaves) {
    const int a = gt_ptr[j], m = gt_ptr[j + 1] - a;
    double sS = 0, sT = 0, s0 = 0;
    for (int p = lane; p < m * m; p += 64) {
      const int ii = p / m, kk = p - ii * m;
      const int i = gt_col[a + ii], k = gt_col[a + kk];
      if (mask && (mask[i] || mask[k])) continue;
      const double w = gt_val[a + ii] * gt_val[a + kk];
      const int pos = find_col(col, rp[i], rp[i + 1], k);
      if (pos >= 0) {
        sS += w * st[pos].x;
        sT += w * st[pos].y;
      }
      if (s0p) {
        const int q = find_col(s0c, s0p[i], s0p[i + 1], k);
        if (q >= 0) s0 += w * s0v[q];
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      sS += __shfl_xor(sS, o);
      sT += __shfl_xor(sT, o);
      s0 += __shfl_xor(s0, o);
    }
    if (lane == 0) {
      out[3 * j] = sS;
      out[3 * j + 1] = sT;
      out[3 * j + 2] = s0;
    }
  }
}

// winv[j] = theta / (cs dS + (ct + i cti) dT + i c0 dS0), 0 where the diagonal vanishes
template <typename C>
__global__ void k_grad_winv(int64_t nlag, const double* d3, double cs, double ct, double cti,
                            double c0, double theta, C* winv) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nlag;
       j += (int64_t)gridDim.x * blockDim.x) {
    const double re = cs * d3[3 * j] + ct * d3[3 * j + 1];
    const double im = cti * d3[3 * j + 1] + c0 * d3[3 * j + 2];
    const double den = re * re + im * im;
    winv[j] = den > 0.0 ? from_d<C>(make_double2(theta * re / den, -theta * im / den))
                        : from_d<C>(make_double2(0.0, 0.0));
  }
}

// ---- application without stored matrices ------------------------------------------
// Every entry of G and Pi follows from the DOF's entity (edge e = (a, b), a < b; face
// f = (p, q, r), p < q < r; DOF layout 2e + k and 2 nE + 2f + k for order 2, e for
// order 1) and the coordinates:
//
//   G, order 2 (units of 1/3):  edge DOF k = 0: -5 w_a + w_b + 4 w_e,  k = 1: -w_a + 5 w_b
//     - 4 w_e; face DOF k = 0: -w_p + w_q - 4 w_(p,r) + 4 w_(q,r), k = 1: -w_p + w_r
//     - 4 w_(p,q) + 4 w_(q,r) (w_e: the P2 node at the midpoint of edge e);
//   G, order 1: w_b - w_a;
//   Pi: edge DOF k: (x_b - x_a) . ((2 - k)/3 u_a + (1 + k)/3 u_b) (order 1: (u_a + u_b)/2);
//     face DOF k: (x_q - x_p or x_r - x_p) . (u_p + u_q + u_r) / 3.
//
// These are the nodal duals (tangential derivative / value at the DOF point along its
// tangent) of the P2 (P1) Lagrange basis and of P1 vector fields; G^T and Pi^T run over
// vertex -> edge / face and edge -> face adjacency lists.  The stored CSR forms of G and
// Pi^T are used only to form diag(G^T A G) and diag(Pi^T A Pi) once per system.

// (G w)_i and (Pi u)_i for DOF i
template <typename R, int ORDER, bool NODAL>
__device__ __forceinline__ void aux_row(int i, int nv, int nE, const int32_t* __restrict__ edges,
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
    const C wa = w[a], wb = w[b];
    if constexpr (ORDER == 2) {
      const C we = w[nv + e];
      const R ca = k ? R(-1) : R(-5), cb = k ? R(5) : R(1), ce = k ? R(-4) : R(4);
      gw = mk(ca * wa.x + cb * wb.x + ce * we.x, ca * wa.y + cb * wb.y + ce * we.y);
    } else {
      gw = mk(wb.x - wa.x, wb.y - wa.y);
    }
    if constexpr (NODAL) {
      const R lb = ORDER == 2 ? R(k + 1) * R(1.0 / 3.0) : R(0.5);
      const R la = R(1) - lb;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const R t = xyz[3 * b + d] - xyz[3 * a + d];
        const C ua = u[3 * a + d], ub = u[3 * b + d];
        pu.x += t * (la * ua.x + lb * ub.x);
        pu.y += t * (la * ua.y + lb * ub.y);
      }
    }
  } else if constexpr (ORDER == 2) {
    const int f = (i - ne_dofs) >> 1;
    const int k = i & 1;
    const int p = faces[3 * f], q = faces[3 * f + 1], r = faces[3 * f + 2];
    const int o = k ? r : q;
    const C wp = w[p], wo = w[o];
    const C we = w[nv + face_edges[3 * f + (k ? 0 : 1)]];   // (p,q) for k = 1, (p,r) for k = 0
    const C wqr = w[nv + face_edges[3 * f + 2]];
    gw = mk(wo.x - wp.x + R(4) * (wqr.x - we.x), wo.y - wp.y + R(4) * (wqr.y - we.y));
    if constexpr (NODAL) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const R t = (xyz[3 * o + d] - xyz[3 * p + d]) * R(1.0 / 3.0);
        const C up = u[3 * p + d], uq = u[3 * q + d], ur = u[3 * r + d];
        pu.x += t * (up.x + uq.x + ur.x);
        pu.y += t * (up.y + uq.y + ur.y);
      }
    }
  }
}

// z = dinv .* r + inv_den G w (+ Pi u) (masked rows zero); optional reduction of r.z and
// |r|^2.
template <typename R, int ORDER, bool NODAL>
__global__ __launch_bounds__(TB) void k_aux_z(int n, int G, int nv, int nE, R inv_den,
                                              const int32_t* __restrict__ edges,
                                              const int32_t* __restrict__ faces,
                                              const int32_t* __restrict__ face_edges,
                                              const R* __restrict__ xyz,
                                              const typename C2<R>::T* __restrict__ w,
                                              const typename C2<R>::T* __restrict__ u,
                                              const typename C2<R>::T* __restrict__ dinv,
                                              const typename C2<R>::T* __restrict__ r,
                                              typename C2<R>::T* __restrict__ z,
                                              const uint8_t* __restrict__ mask, int reduce,
                                              RedSite rs) {
  using C = typename C2<R>::T;
  const int lb = xcd_block(G);
  const int chunk = (n + G - 1) / G;
  const int start = lb * chunk, end = min(n, start + chunk);
  double2 d1 = make_double2(0, 0), d2 = make_double2(0, 0);
  for (int i = start + threadIdx.x; i < end; i += TB) {
    const C ri = r[i];
    C gw, pu;
    aux_row<R, ORDER, NODAL>(i, nv, nE, edges, faces, face_edges, xyz, w, u, gw, pu);
    const C di = cmul(dinv[i], ri);
    C zi = mk(di.x + inv_den * gw.x + pu.x, di.y + inv_den * gw.y + pu.y);
    if (mask && mask[i]) zi = mk(R(0), R(0));
    z[i] = zi;
    if (reduce) {
      const double2 rd = to_d(ri), zd = to_d(zi);
      if (rs.paired) {   // two real systems in the real and imaginary parts
        dot_acc(d1, rd, zd, true);
        sq_acc(d2, rd, true);
      } else {
        d1.x += rd.x * zd.x - rd.y * zd.y;
        d1.y += rd.x * zd.y + rd.y * zd.x;
        d2.x += rd.x * rd.x + rd.y * rd.y;
      }
    }
  }
  if (reduce) reduce_finalize<TB>(rs, G, d1, d2);
}

// Order 2: the two DOFs of an edge or a face read the same entity vertices, w, u and
// coordinates, so one thread forms both rows: entity j < nE is the edge j, otherwise the
// face j - nE, and its rows are 2j and 2j + 1 in both cases (r, dinv and z move as one
// 16-byte (FP32) pair per thread).  Same sums as aux_row for k = 0 and k = 1.
template <typename C>
__device__ __forceinline__ void ld_pair(const C* __restrict__ a, int64_t j, C& v0, C& v1) {
  if constexpr (sizeof(C) == 8) {
    const float4 t = reinterpret_cast<const float4*>(a)[j];
    v0 = mk(t.x, t.y);
    v1 = mk(t.z, t.w);
  } else {
    v0 = a[2 * j];
    v1 = a[2 * j + 1];
  }
}

template <typename R, bool NODAL>
__global__ __launch_bounds__(TB) void k_aux_z2(int n2, int G, int nv, int nE, R inv_den,
                                               const int32_t* __restrict__ edges,
                                               const int32_t* __restrict__ faces,
                                               const int32_t* __restrict__ face_edges,
                                               const R* __restrict__ xyz,
                                               const typename C2<R>::T* __restrict__ w,
                                               const typename C2<R>::T* __restrict__ u,
                                               const typename C2<R>::T* __restrict__ dinv,
                                               const typename C2<R>::T* __restrict__ r,
                                               typename C2<R>::T* __restrict__ z,
                                               const uint8_t* __restrict__ mask, int reduce,
                                               RedSite rs) {
  using C = typename C2<R>::T;
  const int lb = xcd_block(G);
  const int chunk = (n2 + G - 1) / G;
  const int start = lb * chunk, end = min(n2, start + chunk);
  double2 d1 = make_double2(0, 0), d2 = make_double2(0, 0);
  for (int j = start + threadIdx.x; j < end; j += TB) {
    C g0, g1, p0 = mk(R(0), R(0)), p1 = mk(R(0), R(0));
    if (j < nE) {
      const int a = edges[2 * j], b = edges[2 * j + 1];
      const C wa = w[a], wb = w[b], we = w[nv + j];
      g0 = mk(R(-5) * wa.x + wb.x + R(4) * we.x, R(-5) * wa.y + wb.y + R(4) * we.y);
      g1 = mk(R(5) * wb.x - wa.x - R(4) * we.x, R(5) * wb.y - wa.y - R(4) * we.y);
      if constexpr (NODAL) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const R t = xyz[3 * b + d] - xyz[3 * a + d];
          const C ua = u[3 * a + d], ub = u[3 * b + d];
          const R t1 = t * R(1.0 / 3.0), t2 = t * R(2.0 / 3.0);
          p0.x += t2 * ua.x + t1 * ub.x;
          p0.y += t2 * ua.y + t1 * ub.y;
          p1.x += t1 * ua.x + t2 * ub.x;
          p1.y += t1 * ua.y + t2 * ub.y;
        }
      }
    } else {
      const int f = j - nE;
      const int fp = faces[3 * f], fq = faces[3 * f + 1], fr = faces[3 * f + 2];
      const C wp = w[fp], wq = w[fq], wr = w[fr];
      const C wpq = w[nv + face_edges[3 * f]], wpr = w[nv + face_edges[3 * f + 1]];
      const C wqr = w[nv + face_edges[3 * f + 2]];
      g0 = mk(wq.x - wp.x + R(4) * (wqr.x - wpr.x), wq.y - wp.y + R(4) * (wqr.y - wpr.y));
      g1 = mk(wr.x - wp.x + R(4) * (wqr.x - wpq.x), wr.y - wp.y + R(4) * (wqr.y - wpq.y));
      if constexpr (NODAL) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const R xp = xyz[3 * fp + d];
          const R t0 = (xyz[3 * fq + d] - xp) * R(1.0 / 3.0);
          const R t1 = (xyz[3 * fr + d] - xp) * R(1.0 / 3.0);
          const C up = u[3 * fp + d], uq = u[3 * fq + d], ur = u[3 * fr + d];
          const R sx = up.x + uq.x + ur.x, sy = up.y + uq.y + ur.y;
          p0.x += t0 * sx;
          p0.y += t0 * sy;
          p1.x += t1 * sx;
          p1.y += t1 * sy;
        }
      }
    }
    C r0, r1, e0, e1;
    ld_pair(r, j, r0, r1);
    ld_pair(dinv, j, e0, e1);
    const C q0 = cmul(e0, r0), q1 = cmul(e1, r1);
    C z0 = mk(q0.x + inv_den * g0.x + p0.x, q0.y + inv_den * g0.y + p0.y);
    C z1 = mk(q1.x + inv_den * g1.x + p1.x, q1.y + inv_den * g1.y + p1.y);
    if (mask) {
      if (mask[2 * j]) z0 = mk(R(0), R(0));
      if (mask[2 * j + 1]) z1 = mk(R(0), R(0));
    }
    if constexpr (sizeof(C) == 8) {
      reinterpret_cast<float4*>(z)[j] = make_float4(z0.x, z0.y, z1.x, z1.y);
    } else {
      z[2 * j] = z0;
      z[2 * j + 1] = z1;
    }
    if (reduce) {
      const double2 ra = to_d(r0), za = to_d(z0), rb = to_d(r1), zb = to_d(z1);
      if (rs.paired) {   // two real systems in the real and imaginary parts
        d1.x += ra.x * za.x + rb.x * zb.x;
        d1.y += ra.y * za.y + rb.y * zb.y;
        d2.x += ra.x * ra.x + rb.x * rb.x;
        d2.y += ra.y * ra.y + rb.y * rb.y;
      } else {
        d1.x += ra.x * za.x - ra.y * za.y + rb.x * zb.x - rb.y * zb.y;
        d1.y += ra.x * za.y + ra.y * za.x + rb.x * zb.y + rb.y * zb.x;
        d2.x += ra.x * ra.x + ra.y * ra.y + rb.x * rb.x + rb.y * rb.y;
      }
    }
  }
  if (reduce) reduce_finalize<TB>(rs, G, d1, d2);
}

// Vertex rows of G^T and Pi^T, L lanes per vertex v:
//   w_v = winv_v inv_den (G^T r)_v,  u_(v,d) = vinv_(v,d) (Pi^T r)_(v,d)
// over the edges at v (entries (e << 1) | [v = b]) and the faces at v (entries
// (f << 2) | position of v in (p, q, r)).
template <typename R, int L, int ORDER, bool NODAL>
__device__ __forceinline__ void aux_vertex_rows(int64_t blk, int nv, int nE, R inv_den,
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
  const int lane = threadIdx.x % L;
  const int64_t v = (blk * TB + threadIdx.x) / L;
  C g = mk(R(0), R(0));
  C acc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) acc[d] = mk(R(0), R(0));
  if (v < nv) {
    // this vertex's coordinates once: each edge or face then needs only its other vertices'
    R xv[3];
    if constexpr (NODAL) {
#pragma unroll
      for (int d = 0; d < 3; ++d) xv[d] = xyz[3 * v + d];
    }
    for (int q = ve_ptr[v] + lane; q < ve_ptr[v + 1]; q += L) {
      const int ent = ve_ent[q];
      const int e = ent >> 1;
      const bool second = ent & 1;
      C sr;
      if constexpr (ORDER == 2) {
        const C r0 = r[2 * e], r1 = r[2 * e + 1];
        // G: -5 r0 - r1 at a, r0 + 5 r1 at b; Pi: lambda_v = (2/3, 1/3) at a, (1/3, 2/3) at b
        const R g0 = second ? R(1) : R(-5), g1 = second ? R(5) : R(-1);
        g.x += g0 * r0.x + g1 * r1.x;
        g.y += g0 * r0.y + g1 * r1.y;
        const R l0 = second ? R(1.0 / 3.0) : R(2.0 / 3.0), l1 = R(1) - l0;
        sr = mk(l0 * r0.x + l1 * r1.x, l0 * r0.y + l1 * r1.y);
      } else {
        const C r0 = r[e];
        g.x += second ? r0.x : -r0.x;
        g.y += second ? r0.y : -r0.y;
        sr = mk(R(0.5) * r0.x, R(0.5) * r0.y);
      }
      if constexpr (NODAL) {
        const int o = edges[2 * e + (second ? 0 : 1)];   // the edge's other vertex
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const R xo = xyz[3 * o + d];
          const R t = second ? xv[d] - xo : xo - xv[d];   // x_b - x_a
          acc[d].x += t * sr.x;
          acc[d].y += t * sr.y;
        }
      }
    }
    if constexpr (ORDER == 2) {
      for (int q = vf_ptr[v] + lane; q < vf_ptr[v + 1]; q += L) {
        const int ent = vf_ent[q];
        const int f = ent >> 2, pos = ent & 3;
        const C r0 = r[2 * nE + 2 * f], r1 = r[2 * nE + 2 * f + 1];
        // G: -(r0 + r1) at p, r0 at q, r1 at r
        if (pos == 0) { g.x -= r0.x + r1.x; g.y -= r0.y + r1.y; }
        else if (pos == 1) { g.x += r0.x; g.y += r0.y; }
        else { g.x += r1.x; g.y += r1.y; }
        if constexpr (NODAL) {
          // the face's other two vertices, in (p, q, r) order
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
  }
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) {
    g.x += __shfl_xor(g.x, o);
    g.y += __shfl_xor(g.y, o);
  }
  if constexpr (NODAL) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int o = L / 2; o > 0; o >>= 1) {
        acc[d].x += __shfl_xor(acc[d].x, o);
        acc[d].y += __shfl_xor(acc[d].y, o);
      }
  }
  if (lane == 0 && v < nv) {
    w[v] = cmul(winv[v], mk(inv_den * g.x, inv_den * g.y));
    if constexpr (NODAL)
      for (int d = 0; d < 3; ++d) u[3 * v + d] = cmul(vinv[3 * v + d], acc[d]);
  }
}

// Edge-node rows of G^T (order 2), L lanes per edge e: its own two DOFs (4 r0 - 4 r1) and
// the faces containing e (entries (f << 2) | position of e among (p,q), (p,r), (q,r)):
// -4 r1 at (p,q), -4 r0 at (p,r), 4 (r0 + r1) at (q,r); w_(nv+e) = winv inv_den (.).
template <typename R, int L>
__device__ __forceinline__ void aux_edge_rows(int64_t blk, int nv, int nE, R inv_den,
                                               const int32_t* __restrict__ ef_ptr,
                                               const int32_t* __restrict__ ef_ent,
                                               const typename C2<R>::T* __restrict__ r,
                                               const typename C2<R>::T* __restrict__ winv,
                                               typename C2<R>::T* __restrict__ w) {
  using C = typename C2<R>::T;
  const int lane = threadIdx.x % L;
  const int64_t e = (blk * TB + threadIdx.x) / L;
  C g = mk(R(0), R(0));
  if (e < nE) {
    if (lane == 0) {
      const C r0 = r[2 * e], r1 = r[2 * e + 1];
      g = mk(R(4) * (r0.x - r1.x), R(4) * (r0.y - r1.y));
    }
    for (int q = ef_ptr[e] + lane; q < ef_ptr[e + 1]; q += L) {
      const int ent = ef_ent[q];
      const int f = ent >> 2, pos = ent & 3;
      const C r0 = r[2 * nE + 2 * f], r1 = r[2 * nE + 2 * f + 1];
      if (pos == 0) { g.x -= R(4) * r1.x; g.y -= R(4) * r1.y; }
      else if (pos == 1) { g.x -= R(4) * r0.x; g.y -= R(4) * r0.y; }
      else { g.x += R(4) * (r0.x + r1.x); g.y += R(4) * (r0.y + r1.y); }
    }
  }
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) {
    g.x += __shfl_xor(g.x, o);
    g.y += __shfl_xor(g.y, o);
  }
  if (lane == 0 && e < nE) w[nv + e] = cmul(winv[nv + e], mk(inv_den * g.x, inv_den * g.y));
}

// One launch for both kinds of G^T rows: blocks [0, Gv) take the vertex rows (with Pi^T),
// blocks [Gv, gridDim) the edge-node rows, so the two gathers overlap on the device.  Gv
// and gridDim - Gv are multiples of 8 and each part is mapped XCD-aware (each XCD takes a
// contiguous eighth of the vertices and of the edges): neighbouring rows gather the same
// r entries, and only blocks on one XCD share an L2.
template <typename R, int LV, int LE, int ORDER, bool NODAL>
__global__ __launch_bounds__(TB) void k_aux_t(int Gv, int nv, int nE, R inv_den,
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
    aux_vertex_rows<R, LV, ORDER, NODAL>(lb, nv, nE, inv_den, ve_ptr, ve_ent, vf_ptr, vf_ent,
                                         edges, faces, xyz, r, winv, vinv, w, u);
  } else if constexpr (ORDER == 2) {
    const int be = b - Gv, Ge = (int)gridDim.x - Gv;
    const int lb = (be & 7) * (Ge >> 3) + (be >> 3);
    aux_edge_rows<R, LE>(lb, nv, nE, inv_den, ef_ptr, ef_ent, r, winv, w);
  }
}

// ---- vector nodal term --------------------------------------------------------------
// Contribution t = cell * npairs + q, q over the (local DOF i, cell vertex v) pairs with
// lambda_v(x_i) != 0; key (DOF, vertex) or, transposed, (vertex, DOF).
__global__ void k_nodal_keys(int64_t nc, int nloc, int npairs, const int32_t* pairs,
                             const int32_t* cell_dofs, const int32_t* cells, int B, int transpose,
                             uint64_t* keys, uint32_t* vals) {
  const int64_t n = nc * npairs;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / npairs;
    const int q = (int)(t % npairs);
    const uint64_t dof = (uint64_t)cell_dofs[c * nloc + pairs[2 * q]];
    const uint64_t vert = (uint64_t)cells[4 * c + pairs[2 * q + 1]];
    keys[t] = transpose ? ((vert << B) | dof) : ((dof << B) | vert);
    vals[t] = (uint32_t)t;
  }
}

// Value of each unique entry from its first contribution: lambda * (J t_ref)_d, SoA over d.
__global__ void k_nodal_vals(int64_t nnz, const int32_t* seg, const uint32_t* perm, int npairs,
                             const int32_t* pairs, const double* plam, const double* dual_t,
                             const double* xyz, const int32_t* cells, double* v64) {
  for (int64_t z = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; z < nnz;
       z += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t t = perm[seg[z]];
    const int64_t c = t / (uint32_t)npairs;
    const int q = (int)(t % (uint32_t)npairs);
    const int i = pairs[2 * q];
    const double* x0 = xyz + 3 * (int64_t)cells[4 * c];
    const double* x1 = xyz + 3 * (int64_t)cells[4 * c + 1];
    const double* x2 = xyz + 3 * (int64_t)cells[4 * c + 2];
    const double* x3 = xyz + 3 * (int64_t)cells[4 * c + 3];
    const double* tr = dual_t + 3 * i;
    for (int d = 0; d < 3; ++d) {
      const double td = tr[0] * (x1[d] - x0[d]) + tr[1] * (x2[d] - x0[d]) + tr[2] * (x3[d] - x0[d]);
      const double v = plam[q] * td;
      v64[d * nnz + z] = v;
    }
  }
}

// One wavefront per vertex v: out[3 (3v + d) + {0,1,2}] = p^T {S, T, S0} p for the
// column p = (v, d) of Pi over its (unmasked) support.
__global__ __launch_bounds__(TB) void k_nodal_diag(int64_t nv, int64_t nnz, const int32_t* pt_ptr,
                                                   const int32_t* pt_col, const double* pt_val,
                                                   const int32_t* rp, const int32_t* col,
                                                   const double2* st, const int32_t* s0p,
                                                   const int32_t* s0c, const double* s0v,
                                                   const uint8_t* mask, double* out) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t v = wave; v < nv; v += nwaves) {
    const int a = pt_ptr[v], m = pt_ptr[v + 1] - a;
    double sS[3] = {0, 0, 0}, sT[3] = {0, 0, 0}, s0[3] = {0, 0, 0};
    for (int p = lane; p < m * m; p += 64) {
      const int ii = p / m, kk = p - ii * m;
      const int i = pt_col[a + ii], k = pt_col[a + kk];
      if (mask && (mask[i] || mask[k])) continue;
      double w[3];
      for (int d = 0; d < 3; ++d) w[d] = pt_val[d * nnz + a + ii] * pt_val[d * nnz + a + kk];
      const int pos = find_col(col, rp[i], rp[i + 1], k);
      if (pos >= 0)
        for (int d = 0; d < 3; ++d) {
          sS[d] += w[d] * st[pos].x;
          sT[d] += w[d] * st[pos].y;
        }
      if (s0p) {
        const int q = find_col(s0c, s0p[i], s0p[i + 1], k);
        if (q >= 0)
          for (int d = 0; d < 3; ++d) s0[d] += w[d] * s0v[q];
      }
    }
    for (int d = 0; d < 3; ++d)
      for (int o = 32; o > 0; o >>= 1) {
        sS[d] += __shfl_xor(sS[d], o);
        sT[d] += __shfl_xor(sT[d], o);
        s0[d] += __shfl_xor(s0[d], o);
      }
    if (lane == 0)
      for (int d = 0; d < 3; ++d) {
        out[3 * (3 * v + d)] = sS[d];
        out[3 * (3 * v + d) + 1] = sT[d];
        out[3 * (3 * v + d) + 2] = s0[d];
      }
  }
}

// vertex -> entity adjacency entries: edges (v, (e << 1) | role), faces (v, f << 2 | role)
__global__ void k_adj_keys(int64_t nent, int nvper, const int32_t* ent_verts, int B,
                           int shift, uint64_t* keys, uint32_t* vals) {
  const int64_t n = nent * nvper;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = t / nvper;
    const int role = (int)(t % nvper);
    const uint64_t code = ((uint64_t)e << shift) | (uint64_t)role;
    keys[t] = ((uint64_t)ent_verts[t] << B) | code;
    vals[t] = (uint32_t)t;
  }
}

__global__ void k_xyz32(int64_t n, const double* a, float* b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// face_edges[3F + j] = the edges (p,q), (p,r), (q,r) of face F = (p, q, r), from each cell's
// local face f and UFC local edges (every cell that holds F writes the same ids)
__constant__ int c_face_edge[4][3] = {{2, 1, 0}, {4, 3, 0}, {5, 3, 1}, {5, 4, 2}};
__global__ void k_face_edges(int64_t nc, const int32_t* cell_faces, const int32_t* cell_edges,
                             int32_t* face_edges) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nc * 4;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / 4;
    const int f = (int)(t % 4);
    const int64_t F = cell_faces[t];
    for (int j = 0; j < 3; ++j) face_edges[3 * F + j] = cell_edges[6 * c + c_face_edge[f][j]];
  }
}

// edge -> face adjacency entries (e, (f << 2) | position of e among (p,q), (p,r), (q,r))
__global__ void k_ef_keys(int64_t nF, const int32_t* face_edges, int B, uint64_t* keys,
                          uint32_t* vals) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < 3 * nF;
       t += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t code = ((uint64_t)(t / 3) << 2) | (uint64_t)(t % 3);
    keys[t] = ((uint64_t)face_edges[t] << B) | code;
    vals[t] = (uint32_t)t;
  }
}

}  // namespace

void ensure_nodal(Context* c, Space* s) {
  std::lock_guard<std::mutex> hold(s->lazy);
  if (s->pit_ptr.p) return;
  const RefData& R = ref_data(s->order);
  const Mesh* m = s->mesh;
  std::vector<int32_t> pairs;
  std::vector<double> plam;
  for (int i = 0; i < R.nloc; ++i)
    for (int v = 0; v < 4; ++v)
      if (R.dual_lam[4 * i + v] != 0.0) {
        pairs.push_back(i);
        pairs.push_back(v);
        plam.push_back(R.dual_lam[4 * i + v]);
      }
  const int npairs = (int)plam.size();
  const int64_t n = m->nc * npairs;
  SUCEM_CHECK(n < (int64_t(1) << 31), SUCEM_ERR_UNSUPPORTED, "nodal interpolation too large");
  DBuf<int32_t> pairs_d(pairs.size());
  DBuf<double> plam_d(plam.size()), dual_t_d(R.dual_t.size());
  h2d(c, pairs_d.p, pairs.data(), pairs.size());
  h2d(c, plam_d.p, plam.data(), plam.size());
  h2d(c, dual_t_d.p, R.dual_t.data(), R.dual_t.size());
  const int B = key_bits(std::max<int64_t>(s->ndofs, m->nv));
  {
    // Pi^T with stored FP64 values (for diag(Pi^T A Pi) only)
    DBuf<uint64_t> keys(n);
    DBuf<uint32_t> vals(n);
    SUCEM_LAUNCH(k_nodal_keys, grid_for(n, TB), TB, 0, c->stream, m->nc, s->nloc, npairs,
                 pairs_d.p, s->cell_dofs.p, m->cells.p, B, 1, keys.p, vals.p);
    DBuf<int32_t> seg;
    DBuf<uint32_t> perm;
    const int64_t nnz = build_csr_from_pairs(c, n, keys, vals, B, m->nv, s->pit_ptr, s->pit_col,
                                             seg, perm);
    s->pit_val.alloc(3 * nnz);
    SUCEM_LAUNCH(k_nodal_vals, grid_for(nnz, TB), TB, 0, c->stream, nnz, seg.p, perm.p, npairs,
                 pairs_d.p, plam_d.p, dual_t_d.p, m->xyz.p, m->cells.p, s->pit_val.p);
    s->nnz_pi = nnz;
    c->sync();
  }
}

// Vertex -> edge / face and edge -> face adjacency, face -> edge ids and FP32 coordinates:
// what the preconditioner kernels read in place of stored G, G^T, Pi and Pi^T.
void ensure_entity_maps(Context* c, Space* s) {
  std::lock_guard<std::mutex> hold(s->lazy);
  if (s->xyz32.p) return;
  const Mesh* m = s->mesh;
  for (int kind = 0; kind < (s->order == 2 ? 2 : 1); ++kind) {
    const int64_t nent = kind == 0 ? s->nE : s->nF;
    const int nvper = kind == 0 ? 2 : 3;
    const int shift = kind == 0 ? 1 : 2;
    const int Bk = key_bits(std::max<int64_t>(nent << shift, m->nv));
    const int64_t na = nent * nvper;
    DBuf<uint64_t> keys(na);
    DBuf<uint32_t> vals(na);
    SUCEM_LAUNCH(k_adj_keys, grid_for(na, TB), TB, 0, c->stream, nent, nvper,
                 kind == 0 ? s->edges.p : s->faces.p, Bk, shift, keys.p, vals.p);
    DBuf<int32_t> seg;
    DBuf<uint32_t> perm;
    build_csr_from_pairs(c, na, keys, vals, Bk, m->nv, kind == 0 ? s->ve_ptr : s->vf_ptr,
                         kind == 0 ? s->ve_ent : s->vf_ent, seg, perm);
    c->sync();
  }
  if (s->order == 2) {
    s->face_edges.alloc(3 * s->nF);
    SUCEM_LAUNCH(k_face_edges, grid_for(m->nc * 4, TB), TB, 0, c->stream, m->nc,
                 s->cell_faces.p, s->cell_edges.p, s->face_edges.p);
    const int64_t na = 3 * s->nF;
    const int Bk = key_bits(std::max<int64_t>(s->nF << 2, s->nE));
    DBuf<uint64_t> keys(na);
    DBuf<uint32_t> vals(na);
    SUCEM_LAUNCH(k_ef_keys, grid_for(na, TB), TB, 0, c->stream, s->nF, s->face_edges.p, Bk,
                 keys.p, vals.p);
    DBuf<int32_t> seg;
    DBuf<uint32_t> perm;
    build_csr_from_pairs(c, na, keys, vals, Bk, s->nE, s->ef_ptr, s->ef_ent, seg, perm);
    c->sync();
  } else {
    s->vf_ptr.alloc(m->nv + 1);
    SUCEM_HIP(hipMemsetAsync(s->vf_ptr.p, 0, (m->nv + 1) * sizeof(int32_t), c->stream));
    s->vf_ent.alloc(1);
    s->face_edges.alloc(1);
    s->ef_ptr.alloc(s->nE + 1);
    SUCEM_HIP(hipMemsetAsync(s->ef_ptr.p, 0, (s->nE + 1) * sizeof(int32_t), c->stream));
    s->ef_ent.alloc(1);
  }
  s->xyz32.alloc(3 * m->nv);
  SUCEM_LAUNCH(k_xyz32, grid_for(3 * m->nv, TB), TB, 0, c->stream, 3 * m->nv, m->xyz.p,
               s->xyz32.p);
  c->sync();
}

const double* nodal_diag(Context* c, System* sys, bool masked) {
  Space* s = sys->space;
  ensure_nodal(c, s);
  DBuf<double>& d = sys->ndiag[masked ? 1 : 0];
  if (masked && sys->ndiag_pec_version != s->pec_version) d.release();
  if (!d.p) {
    if (masked) sys->ndiag_pec_version = s->pec_version;
    SUCEM_CHECK(sys->st64.p, SUCEM_ERR_INVALID, "S and T must be assembled");
    const int64_t nv = s->mesh->nv;
    d.alloc(9 * nv);
    const bool has_s0 = (sys->which & SUCEM_MAT_S0) && s->nnz0 > 0;
    SUCEM_LAUNCH(k_nodal_diag, grid_for(nv * 64, TB), TB, 0, c->stream, nv, s->nnz_pi,
                 s->pit_ptr.p, s->pit_col.p, s->pit_val.p, s->row_ptr.p, s->col.p, sys->st64.p,
                 has_s0 ? s->s0_ptr.p : nullptr, s->s0_col.p, sys->s0_64.p,
                 masked ? s->pec.p : nullptr, d.p);
    c->sync();
  }
  return d.p;
}

// Weight of the vector nodal term, theta_v Pi diag(Pi^T A Pi)^-1 Pi^T: 2 for driven
// systems (c0 or the preconditioner's complex shift nonzero), 1 for the real shifted
// cavity operator S - sigma T.  Measured on the device, theta_v = 1 / 1.5 / 2 / 2.5 with
// the shifted driven preconditioner (beta = 0.25): 40^3 order-2 cube FP32+IR 1,616 /
// 1,440 / 1,344 / 1,584 COCG iterations, FP64 1,552 / 1,376 / 1,312 / 1,248; 40^3 order-1
// sweep mean 888 / 822 / 788 (FP32+IR) and 764 / 708 / 686 (FP64) at 1 / 1.5 / 2.  The
// oracle agrees (16^3 / 20^3 order-2 driven cube, FP64: 609 / 574 / 571 and 773 / 709 /
// 656 at 1 / 1.5 / 2).  On the cavity operator 2 gains 2 % (60x30x45 order 2, eigen inner
// iterations 54,152 / 53,176 at 1 / 2) but loses 12 % of the serving trace's throughput
// (its eigen requests, sigma at half the lowest mode: 0.48 / 0.42 req/s).
double nodal_weight(const Coef& k) { return (k.s0 != 0.0 || k.ti != 0.0) ? 2.0 : 1.0; }

template <typename R>
void nodal_winv(Context* c, System* sys, Coef k, bool masked, void* vinv) {
  const double* d9 = nodal_diag(c, sys, masked);
  using C = typename C2<R>::T;
  const int64_t nn = 3 * sys->space->mesh->nv;
  SUCEM_LAUNCH((k_grad_winv<C>), grid_for(nn, TB), TB, 0, c->stream, nn, d9, k.s, k.t, k.ti, k.s0,
               nodal_weight(k), (C*)vinv);
}
template void nodal_winv<float>(Context*, System*, Coef, bool, void*);
template void nodal_winv<double>(Context*, System*, Coef, bool, void*);

void ensure_gradient(Context* c, Space* s) {
  std::lock_guard<std::mutex> hold(s->lazy);
  if (s->g_ptr.p) return;
  const RefData& R = ref_data(s->order);
  const Mesh* m = s->mesh;
  s->nlag = s->order == 1 ? m->nv : m->nv + s->nE;
  std::vector<int32_t> pairs, pnum;
  std::vector<double> pvals;
  // local entries are integers (order 1) or thirds (order 2): k / den with |k| <= 7
  s->g_den = s->order == 1 ? 1.0 : 3.0;
  for (int i = 0; i < R.nloc; ++i)
    for (int j = 0; j < R.nlag; ++j)
      if (R.grad[i * R.nlag + j] != 0.0) {
        const double v = R.grad[i * R.nlag + j];
        const long k = std::lround(v * s->g_den);
        SUCEM_CHECK(k != 0 && k >= -8 && k <= 7 && std::fabs(k / s->g_den - v) < 1e-12,
                    SUCEM_ERR_INVALID, "unexpected discrete-gradient entry");
        pairs.push_back(i);
        pairs.push_back(j);
        pvals.push_back(v);
        pnum.push_back((int32_t)k);
      }
  const int npairs = (int)pvals.size();
  SUCEM_CHECK(std::max<int64_t>(s->ndofs, m->nv + s->nE) < (int64_t(1) << 28),
              SUCEM_ERR_UNSUPPORTED, "too many DOFs for the packed gradient");
  DBuf<int32_t> pairs_d(pairs.size()), pnum_d(pnum.size());
  DBuf<double> pvals_d(pvals.size());
  h2d(c, pairs_d.p, pairs.data(), pairs.size());
  h2d(c, pvals_d.p, pvals.data(), pvals.size());
  h2d(c, pnum_d.p, pnum.data(), pnum.size());
  const int64_t n = m->nc * npairs;
  SUCEM_CHECK(n < (int64_t(1) << 31), SUCEM_ERR_UNSUPPORTED, "gradient too large");
  const int B = key_bits(std::max<int64_t>(s->ndofs, s->nlag));
  for (int transpose = 0; transpose < 2; ++transpose) {
    DBuf<uint64_t> keys(n);
    DBuf<uint32_t> vals(n);
    SUCEM_LAUNCH(k_grad_keys, grid_for(n, TB), TB, 0, c->stream, m->nc, s->nloc, R.nlag, npairs,
                 pairs_d.p, s->cell_dofs.p, m->cells.p, s->cell_edges.p, m->nv, B, transpose,
                 keys.p, vals.p);
    DBuf<int32_t> seg;
    DBuf<uint32_t> perm;
    DBuf<int32_t>& ptr = transpose ? s->gt_ptr : s->g_ptr;
    DBuf<int32_t>& colb = transpose ? s->gt_col : s->g_col;
    DBuf<double>& v64 = transpose ? s->gt_val : s->g_val;
    DBuf<float>& v32 = transpose ? s->gt_val32 : s->g_val32;
    DBuf<uint32_t>& pk = transpose ? s->gt_pk : s->g_pk;
    const int64_t nnz = build_csr_from_pairs(c, n, keys, vals, B, transpose ? s->nlag : s->ndofs,
                                             ptr, colb, seg, perm);
    v64.alloc(nnz);
    v32.alloc(nnz);
    pk.alloc(nnz);
    SUCEM_LAUNCH(k_grad_vals, grid_for(nnz, TB), TB, 0, c->stream, nnz, seg.p, perm.p, npairs,
                 pvals_d.p, pnum_d.p, colb.p, v64.p, v32.p, pk.p);
    s->nnz_g = nnz;
    c->sync();
  }
}

const double* gradient_diag(Context* c, System* sys, bool masked) {
  Space* s = sys->space;
  ensure_gradient(c, s);
  DBuf<double>& d = sys->gdiag[masked ? 1 : 0];
  if (masked && sys->gdiag_pec_version != s->pec_version) d.release();   // PEC mask changed
  if (!d.p) {
    if (masked) sys->gdiag_pec_version = s->pec_version;
    SUCEM_CHECK(sys->st64.p, SUCEM_ERR_INVALID, "S and T must be assembled");
    d.alloc(3 * s->nlag);
    const bool has_s0 = (sys->which & SUCEM_MAT_S0) && s->nnz0 > 0;
    const int64_t waves = s->nlag;
    SUCEM_LAUNCH(k_grad_diag, grid_for(waves * 64, TB), TB, 0, c->stream, s->nlag, s->gt_ptr.p,
                 s->gt_col.p, s->gt_val.p, s->row_ptr.p, s->col.p, sys->st64.p,
                 has_s0 ? s->s0_ptr.p : nullptr, s->s0_col.p, sys->s0_64.p,
                 masked ? s->pec.p : nullptr, d.p);
    c->sync();
  }
  return d.p;
}

// Weight of the gradient term, M^-1 = D^-1 + theta G diag(G^T A G)^-1 G^T (+ the nodal
// term): diag(G^T A G) is a mass-like matrix (G^T S G = 0) whose diagonal underestimates
// its inverse's action, and a larger weight on the Jacobi approximation of that inverse
// cuts COCG iterations.  Measured with the oracle on driven cubes (kh fixed): gradient
// form theta = 2 (8^3 order 2: 712 -> 627 iterations against theta = 1); auxiliary form
// theta = 3 at order 2 (12^3: 919 -> 504, 16^3: 1303 -> 708) and 2 at order 1 (16^3:
// 426 -> 383).
double gradient_weight(int order, bool aux) { return aux && order == 2 ? 3.0 : 2.0; }

template <typename R>
void grad_winv(Context* c, System* sys, Coef k, bool masked, bool aux, void* winv) {
  const double* d3 = gradient_diag(c, sys, masked);
  const double theta = gradient_weight(sys->space->order, aux);
  using C = typename C2<R>::T;
  SUCEM_LAUNCH((k_grad_winv<C>), grid_for(sys->space->nlag, TB), TB, 0, c->stream,
               sys->space->nlag, d3, k.s, k.t, k.ti, k.s0, theta, (C*)winv);
}
template void grad_winv<float>(Context*, System*, Coef, bool, bool, void*);
template void grad_winv<double>(Context*, System*, Coef, bool, bool, void*);

// z = M^-1 r; with rs.mode != FIN_NONE the grid also reduces r.z and |r|^2.
template <typename R>
void grad_precond(Context* c, const Space* s, const void* winv, const void* dinv, const void* r,
                  void* w, void* z, const uint8_t* mask, int G, const RedSite& rs,
                  const void* vinv, void* u) {
  using C = typename C2<R>::T;
  SUCEM_CHECK(s->xyz32.p, SUCEM_ERR_INVALID, "preconditioner maps not built");
  const R inv_den = (R)(1.0 / s->g_den);
  const R* xyz;
  if constexpr (sizeof(R) == 4) xyz = s->xyz32.p; else xyz = s->mesh->xyz.p;
  const int nv = (int)s->mesh->nv, nE = (int)s->nE;
  const int reduce = rs.mode != FIN_NONE ? 1 : 0;
  // lanes per vertex row / edge-node row; k_aux_t per FP32 application at 40^3 order 2:
  // (LV, LE) = (16, 4) 37.8 us, (16, 2) 33.8, (16, 1) 33.9, (32, 2) 44.7, (8, 2) 31.5,
  // (4, 2) 35.4, (2, 2) 47.0
  constexpr int LV = 8, LE = 2;
  const int Gv = (int)(((int64_t)nv * LV + 8 * TB - 1) / (8 * TB) * 8);   // multiple of 8
#define VT_ARGS Gv, nv, nE, inv_den, s->ve_ptr.p, s->ve_ent.p, s->vf_ptr.p, s->vf_ent.p, \
                s->ef_ptr.p, s->ef_ent.p, s->edges.p, s->faces.p, xyz, (const C*)r, \
                (const C*)winv, (const C*)vinv, (C*)w, (C*)u
#define Z_ARGS (int)s->ndofs, G, nv, nE, inv_den, s->edges.p, s->faces.p, s->face_edges.p, xyz, \
               (const C*)w, (const C*)u, (const C*)dinv, (const C*)r, (C*)z, mask, reduce, rs
  if (s->order == 2) {
    const int Gt = Gv + (int)(((int64_t)nE * LE + 8 * TB - 1) / (8 * TB) * 8);
    if (vinv) SUCEM_LAUNCH((k_aux_t<R, LV, LE, 2, true>), Gt, TB, 0, c->stream, VT_ARGS);
    else SUCEM_LAUNCH((k_aux_t<R, LV, LE, 2, false>), Gt, TB, 0, c->stream, VT_ARGS);
    const int n2 = (int)(s->ndofs / 2);   // 2 nE + 2 nF rows: one thread per edge or face
#define Z2_ARGS n2, G, nv, nE, inv_den, s->edges.p, s->faces.p, s->face_edges.p, xyz, \
                (const C*)w, (const C*)u, (const C*)dinv, (const C*)r, (C*)z, mask, reduce, rs
    if (vinv) SUCEM_LAUNCH((k_aux_z2<R, true>), G, TB, 0, c->stream, Z2_ARGS);
    else SUCEM_LAUNCH((k_aux_z2<R, false>), G, TB, 0, c->stream, Z2_ARGS);
#undef Z2_ARGS
  } else {
    if (vinv) SUCEM_LAUNCH((k_aux_t<R, LV, LE, 1, true>), Gv, TB, 0, c->stream, VT_ARGS);
    else SUCEM_LAUNCH((k_aux_t<R, LV, LE, 1, false>), Gv, TB, 0, c->stream, VT_ARGS);
    if (vinv) SUCEM_LAUNCH((k_aux_z<R, 1, true>), G, TB, 0, c->stream, Z_ARGS);
    else SUCEM_LAUNCH((k_aux_z<R, 1, false>), G, TB, 0, c->stream, Z_ARGS);
  }
#undef VT_ARGS
#undef Z_ARGS
  finish_reduction(c, rs, G);
}
template void grad_precond<float>(Context*, const Space*, const void*, const void*, const void*,
                                  void*, void*, const uint8_t*, int, const RedSite&, const void*,
                                  void*);
template void grad_precond<double>(Context*, const Space*, const void*, const void*, const void*,
                                   void*, void*, const uint8_t*, int, const RedSite&, const void*,
                                   void*);

}  // namespace sucem
