// Post-processing of a solved field on the device: the power flux through the
// mesh boundary and the near-to-far-field transform (NTFF), replacing the DOLFIN
// functionals of PostProcessing/power_flux.py:52-133,
// PostProcessing/surface_ntff.py:36-155 and PostProcessing/variational_ntff.py:64-131.
//
// Surface terms: one thread per (boundary face, quadrature point) evaluates E and
// curl E from the DOF vector (covariant Piola map of the reference basis) and
// stores the sample; the flux and the far-field potentials are then block sums
// over the samples, one grid row per direction.  Partial sums go to the host in
// block order, so every result is deterministic.
//
// Variational terms: the free-space functional <curl E, curl g> - k0^2 <E, g> over
// the cells that touch the boundary through an edge, with g supported on the
// boundary DOFs.  W_c = (S_e - k0^2 T_e) E_c is formed once per marked cell; each
// direction then needs only g at the boundary DOFs (nodal functionals of the
// testing field evaluated at the DOF points).
#include <cmath>
#include <vector>

#include "geom.hpp"
#include "internal.hpp"
#include "kernels.hpp"

namespace sucem {
namespace {

constexpr int TB = 256;
constexpr int QN = 5;            // collapsed Gauss rule, QN^2 points per face
constexpr int NQ = QN * QN;
constexpr int NTFF_PARTS = 64;   // sample partitions per direction

__constant__ int c_pfaces[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
__constant__ int c_pedges[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};
__constant__ int c_face_edges[4][3] = {{0, 1, 2}, {0, 3, 4}, {1, 3, 5}, {2, 4, 5}};
__constant__ int c_psym[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};

struct Sample {
  double w;          // quadrature weight x 2 x face area
  double p[3];       // physical point
  double n[3];       // outward unit normal
  double2 E[3];      // field
  double2 H[3];      // curl E (divided by mu_r when requested)
};

// Gauss-Legendre nodes and weights on [0, 1] (Newton on P_n).
void gauss01(int n, double* x, double* w) {
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = (1.0 - z) / 2.0;
    w[i] = 1.0 / ((1.0 - z * z) * dp * dp);   // 2/((1-z^2)P'^2) on [-1,1], halved
  }
}

// (s, t, w) of the collapsed rule on the reference triangle (weights sum to 1/2).
std::vector<double> triangle_rule() {
  double x[QN], w[QN];
  gauss01(QN, x, w);
  std::vector<double> r;
  for (int i = 0; i < QN; ++i)
    for (int j = 0; j < QN; ++j) {
      r.push_back(x[i] * (1.0 - x[j]));
      r.push_back(x[j]);
      r.push_back(w[i] * w[j] * (1.0 - x[j]));
    }
  return r;
}

__global__ void k_samples(int64_t nbf, const double* xyz, const int32_t* cells,
                          const int32_t* bf_cell, const int32_t* bf_local,
                          const int32_t* cell_dofs, const double* ref_basis, int nloc,
                          const double* rule, const double2* x, const double* inv_mu,
                          Sample* out) {
  const int64_t total = nbf * NQ;
  for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < total;
       id += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = id / NQ;
    const int q = (int)(id % NQ);
    const int64_t c = bf_cell[f];
    const int lf = bf_local[f];
    const int32_t* cell = cells + 4 * c;
    double J[3][3], Ji[3][3], det;
    cell_jacobian(xyz, cell, J, det, Ji);
    const double* P = xyz + 3 * (int64_t)cell[c_pfaces[lf][0]];
    const double* Q = xyz + 3 * (int64_t)cell[c_pfaces[lf][1]];
    const double* R = xyz + 3 * (int64_t)cell[c_pfaces[lf][2]];
    const double* O = xyz + 3 * (int64_t)cell[lf];   // the vertex opposite the face
    double e1[3], e2[3], n[3];
    for (int k = 0; k < 3; ++k) { e1[k] = Q[k] - P[k]; e2[k] = R[k] - P[k]; }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const double a2 = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; ++k) n[k] /= a2;
    if (n[0] * (O[0] - P[0]) + n[1] * (O[1] - P[1]) + n[2] * (O[2] - P[2]) > 0)
      for (int k = 0; k < 3; ++k) n[k] = -n[k];
    const double s = rule[3 * q], t = rule[3 * q + 1];
    Sample smp;
    smp.w = rule[3 * q + 2] * a2;
    for (int k = 0; k < 3; ++k) {
      smp.p[k] = P[k] + s * e1[k] + t * e2[k];
      smp.n[k] = n[k];
    }
    const double* x0 = xyz + 3 * (int64_t)cell[0];
    double d[3] = {smp.p[0] - x0[0], smp.p[1] - x0[1], smp.p[2] - x0[2]};
    double xh[3];
    for (int a = 0; a < 3; ++a) xh[a] = Ji[a][0] * d[0] + Ji[a][1] * d[1] + Ji[a][2] * d[2];
    const double mono[10] = {1.0,           xh[0],         xh[1],         xh[2],
                             xh[0] * xh[0], xh[0] * xh[1], xh[0] * xh[2], xh[1] * xh[1],
                             xh[1] * xh[2], xh[2] * xh[2]};
    double2 Eh[3], Ch[3];
    for (int k = 0; k < 3; ++k) { Eh[k] = make_double2(0, 0); Ch[k] = make_double2(0, 0); }
    const double* rcurl = ref_basis + nloc * 30;
    for (int i = 0; i < nloc; ++i) {
      const double2 xi = x[cell_dofs[c * nloc + i]];
      for (int dd = 0; dd < 3; ++dd) {
        double bv = 0, cv = 0;
        for (int m = 0; m < 10; ++m) {
          bv += ref_basis[(i * 3 + dd) * 10 + m] * mono[m];
          cv += rcurl[(i * 3 + dd) * 10 + m] * mono[m];
        }
        Eh[dd].x += bv * xi.x; Eh[dd].y += bv * xi.y;
        Ch[dd].x += cv * xi.x; Ch[dd].y += cv * xi.y;
      }
    }
    const double hs = (inv_mu ? inv_mu[c] : 1.0) / det;
    for (int a = 0; a < 3; ++a) {
      // E = J^-T E^, curl E = J C^ / det J
      smp.E[a] = make_double2(Ji[0][a] * Eh[0].x + Ji[1][a] * Eh[1].x + Ji[2][a] * Eh[2].x,
                              Ji[0][a] * Eh[0].y + Ji[1][a] * Eh[1].y + Ji[2][a] * Eh[2].y);
      smp.H[a] = make_double2(hs * (J[a][0] * Ch[0].x + J[a][1] * Ch[1].x + J[a][2] * Ch[2].x),
                              hs * (J[a][0] * Ch[0].y + J[a][1] * Ch[1].y + J[a][2] * Ch[2].y));
    }
    out[id] = smp;
  }
}

// w n . (E_r x (-H_i) + E_i x H_r), summed per block
__global__ __launch_bounds__(TB) void k_flux(int64_t ns, const Sample* smp, double* part) {
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)TB + threadIdx.x; i < ns; i += (int64_t)gridDim.x * TB) {
    const Sample& s = smp[i];
    double a[3] = {s.E[0].x, s.E[1].x, s.E[2].x}, b[3] = {-s.H[0].y, -s.H[1].y, -s.H[2].y};
    double c[3] = {s.E[0].y, s.E[1].y, s.E[2].y}, d[3] = {s.H[0].x, s.H[1].x, s.H[2].x};
    double v = 0;
    for (int k = 0; k < 3; ++k) {
      const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
      v += s.n[k] * (a[k1] * b[k2] - a[k2] * b[k1] + c[k1] * d[k2] - c[k2] * d[k1]);
    }
    acc += s.w * v;
  }
  const double2 r = block_sum<TB>(make_double2(acc, 0.0));
  if (threadIdx.x == 0) part[blockIdx.x] = r.x;
}

__device__ inline void cross_c(const double n[3], const double2 v[3], double2 out[3]) {
  for (int k = 0; k < 3; ++k) {
    const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
    out[k] = make_double2(n[k1] * v[k2].x - n[k2] * v[k1].x, n[k1] * v[k2].y - n[k2] * v[k1].y);
  }
}

// Per direction (grid row y): sum over samples of w exp(j k0 p.r) (n x H) and
// w exp(j k0 p.r) (n x E), projected on theta_hat and phi_hat:
// part[y][x][0..7] = (A.th, A.ph, B.th, B.ph) as complex pairs.
__global__ __launch_bounds__(TB) void k_ntff_surface(int64_t ns, const Sample* smp,
                                                     const double* dirs, double k0,
                                                     double* part) {
  const double* dr = dirs + 9 * blockIdx.y;   // r_hat, theta_hat, phi_hat
  double2 acc[4];
  for (int k = 0; k < 4; ++k) acc[k] = make_double2(0, 0);
  for (int64_t i = blockIdx.x * (int64_t)TB + threadIdx.x; i < ns; i += (int64_t)gridDim.x * TB) {
    const Sample& s = smp[i];
    double sn, cs;
    sincos(k0 * (s.p[0] * dr[0] + s.p[1] * dr[1] + s.p[2] * dr[2]), &sn, &cs);
    double2 nh[3], ne[3];
    cross_c(s.n, s.H, nh);
    cross_c(s.n, s.E, ne);
    for (int v = 0; v < 2; ++v) {
      const double* a = dr + 3 + 3 * v;
      double2 ph = make_double2(0, 0), pe = make_double2(0, 0);
      for (int k = 0; k < 3; ++k) {
        ph.x += a[k] * nh[k].x; ph.y += a[k] * nh[k].y;
        pe.x += a[k] * ne[k].x; pe.y += a[k] * ne[k].y;
      }
      acc[v].x += s.w * (cs * ph.x - sn * ph.y);
      acc[v].y += s.w * (cs * ph.y + sn * ph.x);
      acc[2 + v].x += s.w * (cs * pe.x - sn * pe.y);
      acc[2 + v].y += s.w * (cs * pe.y + sn * pe.x);
    }
  }
  double* out = part + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 8;
  for (int k = 0; k < 4; ++k) {
    const double2 r = block_sum<TB>(acc[k]);
    if (threadIdx.x == 0) { out[2 * k] = r.x; out[2 * k + 1] = r.y; }
  }
}

__global__ void k_mark_edges(int64_t nbf, const int32_t* bf_cell, const int32_t* bf_local,
                             const int32_t* cell_edges, uint8_t* edge_on_b) {
  for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nbf;
       f += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = bf_cell[f];
    for (int k = 0; k < 3; ++k) edge_on_b[cell_edges[6 * c + c_face_edges[bf_local[f]][k]]] = 1;
  }
}

__global__ void k_mark_cells(int64_t nc, const int32_t* cell_edges, const uint8_t* edge_on_b,
                             uint8_t* cell_on_b) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc;
       c += (int64_t)gridDim.x * blockDim.x) {
    uint8_t m = 0;
    for (int k = 0; k < 6; ++k) m |= edge_on_b[cell_edges[6 * c + k]];
    cell_on_b[c] = m;
  }
}

// W[c][i] = sum_j (S_e - k0^2 T_e)_ij x_j for marked cells, eps_r = mu_r = 1.
__global__ void k_cell_apply(int64_t nc, const uint8_t* cell_on_b, const double* xyz,
                             const int32_t* cells, const int32_t* cell_dofs,
                             const double* ref_mass, const double* ref_stiff, int nloc,
                             double k0, const double2* x, double2* W) {
  const int64_t total = nc * nloc;
  for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < total;
       id += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = id / nloc;
    const int i = (int)(id % nloc);
    if (!cell_on_b[c]) continue;
    double J[3][3], Ji[3][3], det;
    cell_jacobian(xyz, cells + 4 * c, J, det, Ji);
    const double adet = fabs(det);
    double ct[6], cs[6];
    for (int k = 0; k < 6; ++k) {
      const int a = c_psym[k][0], b = c_psym[k][1];
      ct[k] = -k0 * k0 * adet * (Ji[a][0] * Ji[b][0] + Ji[a][1] * Ji[b][1] + Ji[a][2] * Ji[b][2]);
      cs[k] = (J[0][a] * J[0][b] + J[1][a] * J[1][b] + J[2][a] * J[2][b]) / adet;
    }
    double2 acc = make_double2(0, 0);
    for (int j = 0; j < nloc; ++j) {
      double kij = 0;
      for (int k = 0; k < 6; ++k)
        kij += cs[k] * ref_stiff[(k * nloc + i) * nloc + j] + ct[k] * ref_mass[(k * nloc + i) * nloc + j];
      const double2 xj = x[cell_dofs[c * nloc + j]];
      acc.x += kij * xj.x;
      acc.y += kij * xj.y;
    }
    W[id] = acc;
  }
}

// Nodal functional of local DOF i: point and (unnormalised) tangent from the cell vertices.
__device__ inline void dof_point(const double* xyz, const int32_t* cell, int order, int i,
                                 double p[3], double t[3]) {
  if (order == 1 || i < 12) {
    const int l = order == 1 ? i : i / 2;
    const double s = order == 1 ? 0.5 : ((i & 1) ? 2.0 / 3.0 : 1.0 / 3.0);
    const double* A = xyz + 3 * (int64_t)cell[c_pedges[l][0]];
    const double* B = xyz + 3 * (int64_t)cell[c_pedges[l][1]];
    for (int k = 0; k < 3; ++k) { t[k] = B[k] - A[k]; p[k] = A[k] + s * t[k]; }
  } else {
    const int f = (i - 12) / 2;
    const double* P = xyz + 3 * (int64_t)cell[c_pfaces[f][0]];
    const double* Q = xyz + 3 * (int64_t)cell[c_pfaces[f][1]];
    const double* R = xyz + 3 * (int64_t)cell[c_pfaces[f][2]];
    const double* T = (i & 1) ? R : Q;
    for (int k = 0; k < 3; ++k) { p[k] = (P[k] + Q[k] + R[k]) / 3.0; t[k] = T[k] - P[k]; }
  }
}

// Per direction: sum over marked cells and boundary DOFs of g_a(i) W[c][i] for the
// testing fields g_a = (j/k0) r x (a x r) exp(j k0 r.x), a = theta_hat, phi_hat.
__global__ __launch_bounds__(TB) void k_ntff_variational(
    int64_t nc, const uint8_t* cell_on_b, const double* xyz, const int32_t* cells,
    const int32_t* cell_dofs, const uint8_t* bdof, int nloc, int order, const double2* W,
    const double* dirs, double k0, double* part) {
  const double* dr = dirs + 9 * blockIdx.y;
  double cf[2][3];
  for (int v = 0; v < 2; ++v) {
    const double* a = dr + 3 + 3 * v;
    double axr[3], r3[3];
    for (int k = 0; k < 3; ++k) {
      const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
      axr[k] = a[k1] * dr[k2] - a[k2] * dr[k1];
    }
    for (int k = 0; k < 3; ++k) {
      const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
      r3[k] = dr[k1] * axr[k2] - dr[k2] * axr[k1];
      cf[v][k] = r3[k] / k0;
    }
  }
  double2 acc[2] = {make_double2(0, 0), make_double2(0, 0)};
  const int64_t total = nc * nloc;
  for (int64_t id = blockIdx.x * (int64_t)TB + threadIdx.x; id < total;
       id += (int64_t)gridDim.x * TB) {
    const int64_t c = id / nloc;
    const int i = (int)(id % nloc);
    if (!cell_on_b[c] || !bdof[cell_dofs[id]]) continue;
    double p[3], t[3];
    dof_point(xyz, cells + 4 * c, order, i, p, t);
    double sn, cs;
    sincos(k0 * (p[0] * dr[0] + p[1] * dr[1] + p[2] * dr[2]), &sn, &cs);
    const double2 w = W[id];
    for (int v = 0; v < 2; ++v) {
      const double ft = cf[v][0] * t[0] + cf[v][1] * t[1] + cf[v][2] * t[2];
      // g = j ft (cs + j sn) = ft (-sn + j cs)
      const double gr = -ft * sn, gi = ft * cs;
      acc[v].x += gr * w.x - gi * w.y;
      acc[v].y += gr * w.y + gi * w.x;
    }
  }
  double* out = part + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 4;
  for (int v = 0; v < 2; ++v) {
    const double2 r = block_sum<TB>(acc[v]);
    if (threadIdx.x == 0) { out[2 * v] = r.x; out[2 * v + 1] = r.y; }
  }
}

// sum over marked cells and boundary DOFs of conj(x_i) W[c][i]
__global__ __launch_bounds__(TB) void k_flux_variational(int64_t nc, const uint8_t* cell_on_b,
                                                         const int32_t* cell_dofs,
                                                         const uint8_t* bdof, int nloc,
                                                         const double2* x, const double2* W,
                                                         double* part) {
  double2 acc = make_double2(0, 0);
  const int64_t total = nc * nloc;
  for (int64_t id = blockIdx.x * (int64_t)TB + threadIdx.x; id < total;
       id += (int64_t)gridDim.x * TB) {
    const int64_t c = id / nloc;
    const int32_t dof = cell_dofs[id];
    if (!cell_on_b[c] || !bdof[dof]) continue;
    const double2 g = make_double2(x[dof].x, -x[dof].y);
    const double2 w = W[id];
    acc.x += g.x * w.x - g.y * w.y;
    acc.y += g.x * w.y + g.y * w.x;
  }
  const double2 r = block_sum<TB>(acc);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = r.x; part[2 * blockIdx.x + 1] = r.y; }
}

void directions(int64_t n, const double* theta, const double* phi, std::vector<double>& d) {
  d.resize(9 * n);
  for (int64_t i = 0; i < n; ++i) {
    const double st = std::sin(theta[i]), ct = std::cos(theta[i]);
    const double sp = std::sin(phi[i]), cp = std::cos(phi[i]);
    const double v[9] = {st * cp, st * sp, ct, ct * cp, ct * sp, -st, -sp, cp, 0.0};
    for (int k = 0; k < 9; ++k) d[9 * i + k] = v[k];
  }
}

void make_samples(Context* c, const Space* s, const void* x, bool use_mu, DBuf<Sample>& smp) {
  const std::vector<double> rule = triangle_rule();
  DBuf<double> drule(rule.size());
  h2d(c, drule.p, rule.data(), rule.size());
  smp.alloc((size_t)s->nbf * NQ);
  const Mesh* m = s->mesh;
  SUCEM_LAUNCH(k_samples, grid_for(s->nbf * NQ, TB, 4096), TB, 0, c->stream, s->nbf,
               m->xyz.p, m->cells.p, s->bf_cell.p, s->bf_local.p, s->cell_dofs.p,
               s->ref_basis.p, s->nloc, drule.p, (const double2*)x,
               use_mu ? m->inv_mu_r.p : nullptr, smp.p);
  c->sync();   // drule is released on return
}

struct Skin {
  DBuf<uint8_t> cell_on_b;
  DBuf<double2> W;
};

void make_skin(Context* c, const Space* s, const void* x, double k0, Skin& sk) {
  const Mesh* m = s->mesh;
  DBuf<uint8_t> edge_on_b(s->nE);
  SUCEM_HIP(hipMemsetAsync(edge_on_b.p, 0, s->nE, c->stream));
  sk.cell_on_b.alloc(m->nc);
  sk.W.alloc((size_t)m->nc * s->nloc);
  SUCEM_LAUNCH(k_mark_edges, grid_for(s->nbf, TB, 4096), TB, 0, c->stream, s->nbf, s->bf_cell.p,
               s->bf_local.p, s->cell_edges.p, edge_on_b.p);
  SUCEM_LAUNCH(k_mark_cells, grid_for(m->nc, TB, 4096), TB, 0, c->stream, m->nc, s->cell_edges.p,
               edge_on_b.p, sk.cell_on_b.p);
  SUCEM_LAUNCH(k_cell_apply, grid_for(m->nc * s->nloc, TB, 8192), TB, 0, c->stream, m->nc,
               sk.cell_on_b.p, m->xyz.p, m->cells.p, s->cell_dofs.p, s->ref_mass.p,
               s->ref_stiff.p, s->nloc, k0, (const double2*)x, sk.W.p);
  c->sync();   // edge_on_b is released on return
}

}  // namespace

double surface_flux(Context* c, const Space* s, const void* x, double k0, bool use_mu) {
  DBuf<Sample> smp;
  make_samples(c, s, x, use_mu, smp);
  const int64_t ns = s->nbf * NQ;
  const unsigned G = grid_for(ns, TB, 1024);
  DBuf<double> part(G);
  SUCEM_LAUNCH(k_flux, G, TB, 0, c->stream, ns, smp.p, part.p);
  std::vector<double> h(G);
  d2h(c, h.data(), part.p, G);
  c->sync();
  double sum = 0;
  for (unsigned i = 0; i < G; ++i) sum += h[i];
  return sum / k0 / kZ0;
}

void variational_flux(Context* c, const Space* s, const void* x, double k0, double out[2]) {
  Skin sk;
  make_skin(c, s, x, k0, sk);
  const int64_t total = s->mesh->nc * s->nloc;
  const unsigned G = grid_for(total, TB, 1024);
  DBuf<double> part(2 * (size_t)G);
  SUCEM_LAUNCH(k_flux_variational, G, TB, 0, c->stream, s->mesh->nc, sk.cell_on_b.p,
               s->cell_dofs.p, s->bdof.p, s->nloc, (const double2*)x, sk.W.p, part.p);
  std::vector<double> h(2 * (size_t)G);
  d2h(c, h.data(), part.p, h.size());
  c->sync();
  double sr = 0, si = 0;
  for (unsigned i = 0; i < G; ++i) { sr += h[2 * i]; si += h[2 * i + 1]; }
  // F = sum g W with g = j conj(x) / (k0 Z0); the flux is conj(F)
  const double fr = -si / (k0 * kZ0), fi = sr / (k0 * kZ0);
  out[0] = fr;
  out[1] = -fi;
}

void ntff(Context* c, const Space* s, const void* x, double k0, bool variational, int64_t n,
          const double* theta, const double* phi, double* E_ff) {
  std::vector<double> dirs;
  directions(n, theta, phi, dirs);
  DBuf<double> ddirs(dirs.size());
  h2d(c, ddirs.p, dirs.data(), dirs.size());
  DBuf<Sample> smp;
  make_samples(c, s, x, false, smp);
  const int64_t ns = s->nbf * NQ;
  const int P = NTFF_PARTS;
  DBuf<double> part((size_t)n * P * 8);
  SUCEM_LAUNCH(k_ntff_surface, dim3(P, (unsigned)n), TB, 0, c->stream, ns, smp.p, ddirs.p, k0,
               part.p);
  std::vector<double> hs((size_t)n * P * 8), hv;
  d2h(c, hs.data(), part.p, hs.size());
  if (variational) {
    Skin sk;
    make_skin(c, s, x, k0, sk);
    DBuf<double> pv((size_t)n * P * 4);
    SUCEM_LAUNCH(k_ntff_variational, dim3(P, (unsigned)n), TB, 0, c->stream, s->mesh->nc,
                 sk.cell_on_b.p, s->mesh->xyz.p, s->mesh->cells.p, s->cell_dofs.p, s->bdof.p,
                 s->nloc, s->order, sk.W.p, ddirs.p, k0, pv.p);
    hv.resize((size_t)n * P * 4);
    d2h(c, hv.data(), pv.p, hv.size());
    c->sync();
  }
  c->sync();
  // far-field factor j k0 exp(-j k0) / (4 pi)
  const double fr = k0 * std::sin(k0) / (4 * M_PI), fi = k0 * std::cos(k0) / (4 * M_PI);
  auto mul = [](double ar, double ai, double br, double bi, double& cr, double& ci) {
    cr = ar * br - ai * bi;
    ci = ar * bi + ai * br;
  };
  for (int64_t a = 0; a < n; ++a) {
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int p = 0; p < P; ++p)
      for (int k = 0; k < 8; ++k) acc[k] += hs[((size_t)a * P + p) * 8 + k];
    // N = j/(k0 Z0) A, L = -B (A = int e n x curl E, B = int e n x E)
    const double sN = 1.0 / (k0 * kZ0);
    const double Nth_r = -sN * acc[1], Nth_i = sN * acc[0];
    const double Nph_r = -sN * acc[3], Nph_i = sN * acc[2];
    const double Lth_r = -acc[4], Lth_i = -acc[5], Lph_r = -acc[6], Lph_i = -acc[7];
    double t_r, t_i, p_r, p_i;
    if (!variational) {
      // E_theta = -f (L_phi + Z0 N_theta), E_phi = f (L_theta - Z0 N_phi)
      mul(fr, fi, -(Lph_r + kZ0 * Nth_r), -(Lph_i + kZ0 * Nth_i), t_r, t_i);
      mul(fr, fi, Lth_r - kZ0 * Nph_r, Lth_i - kZ0 * Nph_i, p_r, p_i);
    } else {
      double v[4] = {0, 0, 0, 0};
      for (int p = 0; p < P; ++p)
        for (int k = 0; k < 4; ++k) v[k] += hv[((size_t)a * P + p) * 4 + k];
      // E_theta = f (-L_phi + E_H_theta), E_phi = f (L_theta + E_H_phi)
      mul(fr, fi, -Lph_r + v[0], -Lph_i + v[1], t_r, t_i);
      mul(fr, fi, Lth_r + v[2], Lth_i + v[3], p_r, p_i);
    }
    E_ff[4 * a] = t_r;
    E_ff[4 * a + 1] = t_i;
    E_ff[4 * a + 2] = p_r;
    E_ff[4 * a + 3] = p_i;
  }
}

}  // namespace sucem
