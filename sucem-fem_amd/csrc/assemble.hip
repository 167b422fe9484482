// Assembly of S (curl-curl), T (eps_r-weighted mass) and S0 (first-order ABC)
// into CSR, replacing dolfin.assemble at SystemMatrices.py:54 for the forms of
// Forms.py:71-85 and BoundaryConditions/ABC.py:56-57.
//
// Two kernels per matrix family:
//  1. per cell (or boundary face): geometry in registers -> six symmetric
//     geometric coefficients (J^-1 J^-T, J^T J or the face projector);
//  2. per CSR nonzero: an atomic-free gather over the sorted (row, col) list of
//     element contributions, contracting the six coefficients of each
//     contributing cell with the exact reference tensors held in LDS.
// The gather sums contributions in cell order, so the result is deterministic.
#include "internal.hpp"
#include "geom.hpp"
#include "kernels.hpp"

namespace sucem {
namespace {

constexpr int TB = 256;

__constant__ int c_sym[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
__constant__ int c_faces[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

// coef[c][0..5]  = eps_r |det J| (J^-1 J^-T)_k ; coef[c][6..11] = (1/mu_r)/|det J| (J^T J)_k
__global__ void k_cell_coefs(int64_t nc, const double* xyz, const int32_t* cells,
                             const double* eps, const double* imu, double* coef) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc;
       c += (int64_t)gridDim.x * blockDim.x) {
    double J[3][3], Ji[3][3], det;
    cell_jacobian(xyz, cells + 4 * c, J, det, Ji);
    double adet = fabs(det);
    double e = eps[c] * adet, m = imu[c] / adet;
    for (int k = 0; k < 6; ++k) {
      int a = c_sym[k][0], b = c_sym[k][1];
      double g = Ji[a][0] * Ji[b][0] + Ji[a][1] * Ji[b][1] + Ji[a][2] * Ji[b][2];
      double q = J[0][a] * J[0][b] + J[1][a] * J[1][b] + J[2][a] * J[2][b];
      coef[12 * c + k] = e * g;
      coef[12 * c + 6 + k] = m * q;
    }
  }
}

// H[f][k] = (area_phys / area_ref) (J^-1 (I - n n^T) J^-T)_k for boundary face f
__global__ void k_face_coefs(int64_t nbf, const double* xyz, const int32_t* cells,
                             const int32_t* bf_cell, const int32_t* bf_local, double* H) {
  for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nbf;
       f += (int64_t)gridDim.x * blockDim.x) {
    const int32_t* cell = cells + 4 * (int64_t)bf_cell[f];
    int lf = bf_local[f];
    double J[3][3], Ji[3][3], det;
    cell_jacobian(xyz, cell, J, det, Ji);
    const double* p = xyz + 3 * (int64_t)cell[c_faces[lf][0]];
    const double* q = xyz + 3 * (int64_t)cell[c_faces[lf][1]];
    const double* r = xyz + 3 * (int64_t)cell[c_faces[lf][2]];
    double e1[3], e2[3], n[3];
    for (int k = 0; k < 3; ++k) { e1[k] = q[k] - p[k]; e2[k] = r[k] - p[k]; }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double a2 = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; ++k) n[k] /= a2;
    // reference face areas: sqrt(3)/2 for face 0, 1/2 otherwise
    double aref2 = lf == 0 ? 1.7320508075688772 : 1.0;
    double scale = a2 / aref2;
    // P = I - n n^T ; H = Ji P Ji^T
    double P[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) P[a][b] = (a == b ? 1.0 : 0.0) - n[a] * n[b];
    for (int k = 0; k < 6; ++k) {
      int a = c_sym[k][0], b = c_sym[k][1];
      double s = 0;
      for (int u = 0; u < 3; ++u)
        for (int v = 0; v < 3; ++v) s += Ji[a][u] * P[u][v] * Ji[b][v];
      H[6 * f + k] = scale * s;
    }
  }
}

template <int NLOC>
__global__ __launch_bounds__(TB) void k_gather_volume(int64_t nnz, const int32_t* __restrict__ seg,
                                                      const uint32_t* __restrict__ perm,
                                                      const double* __restrict__ coef,
                                                      const double* __restrict__ ref_mass,
                                                      const double* __restrict__ ref_stiff,
                                                      double2* st64, float2* st32) {
  constexpr int KK = NLOC * NLOC;
  __shared__ double sm[6 * KK];
  __shared__ double ss[6 * KK];
  for (int i = threadIdx.x; i < 6 * KK; i += blockDim.x) {
    sm[i] = ref_mass[i];
    ss[i] = ref_stiff[i];
  }
  __syncthreads();
  for (int64_t z = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; z < nnz;
       z += (int64_t)gridDim.x * blockDim.x) {
    int a = seg[z], b = seg[z + 1];
    double s = 0.0, t = 0.0;
    for (int q = a; q < b; ++q) {
      uint32_t id = perm[q];
      uint32_t cell = id / KK, loc = id - cell * KK;
      const double* cf = coef + 12 * (int64_t)cell;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        t = fma(cf[k], sm[k * KK + loc], t);
        s = fma(cf[6 + k], ss[k * KK + loc], s);
      }
    }
    st64[z] = make_double2(s, t);
    st32[z] = make_float2((float)s, (float)t);
  }
}

template <int NFD>
__global__ __launch_bounds__(TB) void k_gather_face(int64_t nnz0, const int32_t* __restrict__ seg,
                                                    const uint32_t* __restrict__ perm,
                                                    const double* __restrict__ H,
                                                    const int32_t* __restrict__ bf_local,
                                                    const double* __restrict__ ref_face,
                                                    double* s0_64, float* s0_32) {
  constexpr int KK = NFD * NFD;
  __shared__ double sf[4 * 6 * KK];
  for (int i = threadIdx.x; i < 4 * 6 * KK; i += blockDim.x) sf[i] = ref_face[i];
  __syncthreads();
  for (int64_t z = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; z < nnz0;
       z += (int64_t)gridDim.x * blockDim.x) {
    int a = seg[z], b = seg[z + 1];
    double s = 0.0;
    for (int q = a; q < b; ++q) {
      uint32_t id = perm[q];
      uint32_t f = id / KK, loc = id - f * KK;
      int lf = bf_local[f];
      const double* h = H + 6 * (int64_t)f;
#pragma unroll
      for (int k = 0; k < 6; ++k) s = fma(h[k], sf[(lf * 6 + k) * KK + loc], s);
    }
    s0_64[z] = s;
    s0_32[z] = (float)s;
  }
}

__global__ void k_export_st(int64_t nnz, const double2* st, int which, double* out) {
  for (int64_t z = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; z < nnz;
       z += (int64_t)gridDim.x * blockDim.x)
    out[z] = which == 0 ? st[z].x : st[z].y;
}

__global__ void k_zero(int64_t n, double* out) {
  for (int64_t z = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; z < n;
       z += (int64_t)gridDim.x * blockDim.x)
    out[z] = 0.0;
}

// Scatter the compact ABC matrix onto the full pattern (explicit zeros elsewhere).
__global__ void k_export_s0(int64_t n, const int32_t* rp, const int32_t* col, const int32_t* s0p,
                            const int32_t* s0c, const double* s0v, double* out) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    for (int j = s0p[r]; j < s0p[r + 1]; ++j) {
      const int pos = find_col(col, rp[r], rp[r + 1], s0c[j]);
      if (pos >= 0) out[pos] = s0v[j];
    }
  }
}

// A = S - k0^2 T + i k0 S0 on the full pattern (complex128)
__global__ void k_export_lhs(int64_t n, const int32_t* rp, const int32_t* col, const double2* st,
                             const int32_t* s0p, const int32_t* s0c, const double* s0v,
                             double k0, double2* out) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    for (int j = rp[r]; j < rp[r + 1]; ++j)
      out[j] = make_double2(st[j].x - k0 * k0 * st[j].y, 0.0);
    if (!s0p) continue;
    for (int j = s0p[r]; j < s0p[r + 1]; ++j) {
      const int pos = find_col(col, rp[r], rp[r + 1], s0c[j]);
      if (pos >= 0) out[pos].y = k0 * s0v[j];
    }
  }
}

unsigned persistent_grid(Context* c, int64_t n) {
  int64_t cap = (int64_t)c->num_cus * 8;
  return grid_for(n, TB, cap);
}

}  // namespace

void assemble_system(Context* c, Space* s, System* sys, unsigned which) {
  Mesh* m = s->mesh;
  SUCEM_CHECK(m->eps_r.p && m->inv_mu_r.p, SUCEM_ERR_INVALID, "materials not set on the mesh");
  sys->ctx = c;
  sys->space = s;
  sys->which = which;
  if (which & (SUCEM_MAT_S | SUCEM_MAT_T)) {
    DBuf<double> coef(m->nc * 12);
    SUCEM_LAUNCH(k_cell_coefs, grid_for(m->nc, TB), TB, 0, c->stream, m->nc, m->xyz.p, m->cells.p,
                 m->eps_r.p, m->inv_mu_r.p, coef.p);
    sys->st64.alloc(s->nnz);
    sys->st32.alloc(s->nnz);
    if (s->nloc == 6)
      SUCEM_LAUNCH((k_gather_volume<6>), persistent_grid(c, s->nnz), TB, 0, c->stream, s->nnz,
                   s->seg.p, s->perm.p, coef.p, s->ref_mass.p, s->ref_stiff.p, sys->st64.p,
                   sys->st32.p);
    else
      SUCEM_LAUNCH((k_gather_volume<20>), persistent_grid(c, s->nnz), TB, 0, c->stream, s->nnz,
                   s->seg.p, s->perm.p, coef.p, s->ref_mass.p, s->ref_stiff.p, sys->st64.p,
                   sys->st32.p);
    c->sync();
  }
  if (which & SUCEM_MAT_S0) {
    sys->s0_64.alloc(s->nnz0);
    sys->s0_32.alloc(s->nnz0);
    if (s->nnz0) {
      DBuf<double> H(s->nbf * 6);
      SUCEM_LAUNCH(k_face_coefs, grid_for(s->nbf, TB), TB, 0, c->stream, s->nbf, m->xyz.p,
                   m->cells.p, s->bf_cell.p, s->bf_local.p, H.p);
      if (s->nfd == 3)
        SUCEM_LAUNCH((k_gather_face<3>), persistent_grid(c, s->nnz0), TB, 0, c->stream, s->nnz0,
                     s->s0_seg.p, s->s0_perm.p, H.p, s->bf_local.p, s->ref_face.p, sys->s0_64.p,
                     sys->s0_32.p);
      else
        SUCEM_LAUNCH((k_gather_face<8>), persistent_grid(c, s->nnz0), TB, 0, c->stream, s->nnz0,
                     s->s0_seg.p, s->s0_perm.p, H.p, s->bf_local.p, s->ref_face.p, sys->s0_64.p,
                     sys->s0_32.p);
      c->sync();
    }
  }
}

void face_coefs(Context* c, const Space* s, double* H_d) {
  const Mesh* m = s->mesh;
  if (s->nbf)
    SUCEM_LAUNCH(k_face_coefs, grid_for(s->nbf, TB), TB, 0, c->stream, s->nbf, m->xyz.p,
                 m->cells.p, s->bf_cell.p, s->bf_local.p, H_d);
}

void export_lhs(Context* c, const System* sys, double k0, double2* out_d) {
  const Space* s = sys->space;
  SUCEM_CHECK(sys->st64.p, SUCEM_ERR_INVALID, "S/T were not assembled");
  const bool has_s0 = (sys->which & SUCEM_MAT_S0) && s->nnz0 > 0;
  SUCEM_LAUNCH(k_export_lhs, grid_for(s->ndofs, TB), TB, 0, c->stream, s->ndofs, s->row_ptr.p,
               s->col.p, sys->st64.p, has_s0 ? s->s0_ptr.p : nullptr, s->s0_col.p, sys->s0_64.p,
               k0, out_d);
}

void export_values(Context* c, const System* sys, unsigned which_one, double* out_d) {
  const Space* s = sys->space;
  if (which_one == SUCEM_MAT_S || which_one == SUCEM_MAT_T) {
    SUCEM_CHECK(sys->st64.p, SUCEM_ERR_INVALID, "S/T were not assembled");
    SUCEM_LAUNCH(k_export_st, grid_for(s->nnz, TB), TB, 0, c->stream, s->nnz, sys->st64.p,
                 which_one == SUCEM_MAT_S ? 0 : 1, out_d);
  } else if (which_one == SUCEM_MAT_S0) {
    SUCEM_CHECK(sys->which & SUCEM_MAT_S0, SUCEM_ERR_INVALID, "S0 was not assembled");
    SUCEM_LAUNCH(k_zero, grid_for(s->nnz, TB), TB, 0, c->stream, s->nnz, out_d);
    if (s->nnz0)
      SUCEM_LAUNCH(k_export_s0, grid_for(s->ndofs, TB), TB, 0, c->stream, s->ndofs, s->row_ptr.p,
                   s->col.p, s->s0_ptr.p, s->s0_col.p, sys->s0_64.p, out_d);
  } else {
    throw Error(SUCEM_ERR_INVALID, "export expects exactly one of S, T, S0");
  }
}

}  // namespace sucem
