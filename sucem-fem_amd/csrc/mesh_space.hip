// Mesh generation, materials, entity numbering, DOF map and CSR pattern on the device.
//
// Numbering restates DOLFIN's first-encounter rule (the topology computed behind
// dolfin.FunctionSpace at ProblemConfigurations/EMProblem.py:100-105): scanning
// cells in order and local entities in UFC order, a new entity gets the next id.
// The parallel equivalent used here: sort (entity key, 6*cell+local) pairs,
// take each entity's first occurrence (the stable sort keeps the smallest
// occurrence first) and rank entities by it.
#include <string.h>

#include <cmath>
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "internal.hpp"
#include "kernels.hpp"
#include "refdata.hpp"

namespace sucem {
namespace {

__constant__ int c_edges[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};
__constant__ int c_faces[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

int bits_for(int64_t n) {
  int b = 1;
  while ((int64_t(1) << b) < n) ++b;
  return b;
}

// ---- rocPRIM wrappers (temporary storage allocated per call) ----------------
void sort_pairs(Context* c, const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                uint32_t* vout, size_t n, unsigned end_bit) {
  size_t bytes = 0;
  SUCEM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, 0u, end_bit,
                                      c->stream));
  DBuf<char> tmp(bytes ? bytes : 1);
  SUCEM_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, kin, kout, vin, vout, n, 0u, end_bit,
                                      c->stream));
}

void inclusive_scan_i32(Context* c, const int32_t* in, int32_t* out, size_t n) {
  size_t bytes = 0;
  SUCEM_HIP(rocprim::inclusive_scan(nullptr, bytes, in, out, n, rocprim::plus<int32_t>(),
                                    c->stream));
  DBuf<char> tmp(bytes ? bytes : 1);
  SUCEM_HIP(rocprim::inclusive_scan(tmp.p, bytes, in, out, n, rocprim::plus<int32_t>(),
                                    c->stream));
}

void exclusive_scan_i32(Context* c, const int32_t* in, int32_t* out, size_t n) {
  size_t bytes = 0;
  SUCEM_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, int32_t(0), n,
                                    rocprim::plus<int32_t>(), c->stream));
  DBuf<char> tmp(bytes ? bytes : 1);
  SUCEM_HIP(rocprim::exclusive_scan(tmp.p, bytes, in, out, int32_t(0), n,
                                    rocprim::plus<int32_t>(), c->stream));
}

template <typename T>
T read_scalar(Context* c, const T* d) {
  T h;
  SUCEM_HIP(hipMemcpyAsync(&h, d, sizeof(T), hipMemcpyDeviceToHost, c->stream));
  c->sync();
  return h;
}

// ---- kernels -----------------------------------------------------------------
__device__ inline void sort4(int32_t v[4]) {
#define SW(a, b) if (v[a] > v[b]) { int32_t t = v[a]; v[a] = v[b]; v[b] = t; }
  SW(0, 1) SW(2, 3) SW(0, 2) SW(1, 3) SW(1, 2)
#undef SW
}

__global__ void k_box_vertices(int64_t nx, int64_t ny, int64_t nz, double sx, double sy,
                               double sz, double ox, double oy, double oz, double* xyz) {
  int64_t nv = (nx + 1) * (ny + 1) * (nz + 1);
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
       v += (int64_t)gridDim.x * blockDim.x) {
    int64_t ix = v % (nx + 1), t = v / (nx + 1);
    int64_t iy = t % (ny + 1), iz = t / (ny + 1);
    // DOLFIN UnitCube: x = ix/nx, then scaled and shifted (MeshGenerators.py:35-36)
    xyz[3 * v + 0] = (double)ix / (double)nx * sx + ox;
    xyz[3 * v + 1] = (double)iy / (double)ny * sy + oy;
    xyz[3 * v + 2] = (double)iz / (double)nz * sz + oz;
  }
}

__global__ void k_box_cells(int64_t nx, int64_t ny, int64_t nz, int32_t* cells) {
  int64_t nb = nx * ny * nz;
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nb;
       b += (int64_t)gridDim.x * blockDim.x) {
    int64_t bx = b % nx, t = b / nx;
    int64_t by = t % ny, bz = t / ny;
    int64_t sx = nx + 1, sxy = (nx + 1) * (ny + 1);
    int32_t v0 = (int32_t)(bz * sxy + by * sx + bx);
    int32_t v1 = v0 + 1, v2 = v0 + (int32_t)sx, v3 = v1 + (int32_t)sx;
    int32_t v4 = v0 + (int32_t)sxy, v5 = v1 + (int32_t)sxy, v6 = v2 + (int32_t)sxy,
            v7 = v3 + (int32_t)sxy;
    // Kuhn split along the v0-v7 diagonal (DOLFIN UnitCube cell order)
    int32_t six[6][4] = {{v0, v1, v3, v7}, {v0, v1, v7, v5}, {v0, v5, v7, v4},
                         {v0, v3, v2, v7}, {v0, v6, v4, v7}, {v0, v2, v6, v7}};
    for (int k = 0; k < 6; ++k) {
      sort4(six[k]);
      for (int q = 0; q < 4; ++q) cells[(6 * b + k) * 4 + q] = six[k][q];
    }
  }
}

__global__ void k_sort_cells(int64_t nc, int32_t* cells) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc;
       c += (int64_t)gridDim.x * blockDim.x) {
    int32_t v[4] = {cells[4 * c], cells[4 * c + 1], cells[4 * c + 2], cells[4 * c + 3]};
    sort4(v);
    for (int q = 0; q < 4; ++q) cells[4 * c + q] = v[q];
  }
}

__global__ void k_translate(int64_t nv, double* xyz, double sx, double sy, double sz) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
       v += (int64_t)gridDim.x * blockDim.x) {
    xyz[3 * v] -= sx;
    xyz[3 * v + 1] -= sy;
    xyz[3 * v + 2] -= sz;
  }
}

__global__ void k_materials(int64_t nc, const int32_t* region, int nreg, const int64_t* ids,
                            const double* eps, const double* mu, double* eps_c, double* imu_c,
                            unsigned long long* first_bad) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc;
       c += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = region ? region[c] : 0;
    int k = -1;
    for (int i = 0; i < nreg; ++i)
      if (ids[i] == r) { k = i; break; }
    if (k < 0) {
      atomicMin(first_bad, (unsigned long long)c);
      eps_c[c] = 1.0;
      imu_c[c] = 1.0;
    } else {
      eps_c[c] = eps[k];
      imu_c[c] = 1.0 / mu[k];
    }
  }
}

template <int NV, int NL>
__global__ void k_entity_keys(int64_t nc, const int32_t* cells, int B, uint64_t* keys,
                              uint32_t* vals) {
  int64_t n = nc * NL;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t c = i / NL;
    int l = (int)(i % NL);
    uint64_t k = 0;
    for (int q = 0; q < NV; ++q) {
      int lv = NV == 2 ? c_edges[l][q] : c_faces[l][q];
      k = (k << B) | (uint64_t)cells[4 * c + lv];
    }
    keys[i] = k;
    vals[i] = (uint32_t)i;
  }
}

__global__ void k_heads(int64_t n, const uint64_t* keys, int32_t* flags) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// Mark first occurrences by their (cell*NL + local) position.
__global__ void k_mark_first(int64_t n, const int32_t* flags, const uint32_t* vals,
                             int32_t* mark) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (flags[i]) mark[vals[i]] = 1;
}

// Assign ranked entity ids to every (cell, local) occurrence; record entity vertices
// and, for faces, segment starts (used to find boundary faces).
template <int NV>
__global__ void k_assign(int64_t n, const uint64_t* keys, const int32_t* flags,
                         const int32_t* segid, const uint32_t* vals, const int32_t* rank_at,
                         int B, int32_t* cell_ent, int32_t* ent_verts, int32_t* seg_start,
                         int32_t* ent_of_seg) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    // the head of this segment holds the smallest occurrence; its rank is the id
    if (flags[i]) {
      int32_t e = rank_at[vals[i]];
      int32_t s = segid[i] - 1;
      ent_of_seg[s] = e;
      seg_start[s] = (int32_t)i;
      uint64_t k = keys[i];
      uint64_t m = (uint64_t(1) << B) - 1;
      for (int q = NV - 1; q >= 0; --q) {
        ent_verts[e * NV + q] = (int32_t)(k & m);
        k >>= B;
      }
    }
  }
}

__global__ void k_scatter_ids(int64_t n, const int32_t* segid, const uint32_t* vals,
                              const int32_t* ent_of_seg, int32_t* cell_ent) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    cell_ent[vals[i]] = ent_of_seg[segid[i] - 1];
}

// Boundary faces: segments of length one.  isb / bcell / blocal indexed by face id.
__global__ void k_face_boundary(int64_t nseg, int64_t n, const int32_t* seg_start,
                                const int32_t* ent_of_seg, const uint32_t* vals, int32_t* isb,
                                int32_t* bcell, int32_t* blocal, int32_t* bad) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nseg;
       s += (int64_t)gridDim.x * blockDim.x) {
    int64_t a = seg_start[s];
    int64_t b = s + 1 < nseg ? seg_start[s + 1] : n;
    int32_t f = ent_of_seg[s];
    isb[f] = (b - a == 1) ? 1 : 0;
    if (b - a > 2) atomicMax(bad, 1);
    uint32_t v = vals[a];
    bcell[f] = (int32_t)(v / 4);
    blocal[f] = (int32_t)(v % 4);
  }
}

__global__ void k_compact_bfaces(int64_t nF, const int32_t* isb, const int32_t* pos,
                                 const int32_t* bcell, const int32_t* blocal, int32_t* bf_cell,
                                 int32_t* bf_local) {
  for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nF;
       f += (int64_t)gridDim.x * blockDim.x)
    if (isb[f]) {
      bf_cell[pos[f]] = bcell[f];
      bf_local[pos[f]] = blocal[f];
    }
}

__global__ void k_dofmap(int64_t nc, int order, int64_t nE, const int32_t* ce, const int32_t* cf,
                         int32_t* cd) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc;
       c += (int64_t)gridDim.x * blockDim.x) {
    if (order == 1) {
      for (int l = 0; l < 6; ++l) cd[6 * c + l] = ce[6 * c + l];
    } else {
      for (int l = 0; l < 6; ++l) {
        cd[20 * c + 2 * l] = 2 * ce[6 * c + l];
        cd[20 * c + 2 * l + 1] = 2 * ce[6 * c + l] + 1;
      }
      for (int f = 0; f < 4; ++f) {
        cd[20 * c + 12 + 2 * f] = (int32_t)(2 * nE) + 2 * cf[4 * c + f];
        cd[20 * c + 13 + 2 * f] = (int32_t)(2 * nE) + 2 * cf[4 * c + f] + 1;
      }
    }
  }
}

__global__ void k_pec(int64_t nbf, int nloc, int nfd, const int32_t* bf_cell,
                      const int32_t* bf_local, const int32_t* cd, const int32_t* face_dofs,
                      uint8_t* pec) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nbf * nfd;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t f = i / nfd;
    int a = (int)(i % nfd);
    int c = bf_cell[f], lf = bf_local[f];
    pec[cd[(int64_t)c * nloc + face_dofs[lf * nfd + a]]] = 1;
  }
}

// Closure of the flagged facets: every (cell, local face) whose global face is flagged.
__global__ void k_pec_facets(int64_t nc, int nloc, int nfd, const int32_t* cell_faces,
                             const uint8_t* flag, const int32_t* cd, const int32_t* face_dofs,
                             uint8_t* pec) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nc * 4;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (!flag[cell_faces[i]]) continue;
    const int64_t c = i / 4;
    const int lf = (int)(i % 4);
    for (int a = 0; a < nfd; ++a) pec[cd[c * nloc + face_dofs[lf * nfd + a]]] = 1;
  }
}

__global__ void k_cell_pair_keys(int64_t nc, int nloc, int B, const int32_t* cd, uint64_t* keys,
                                 uint32_t* vals) {
  int64_t kk = (int64_t)nloc * nloc;
  int64_t n = nc * kk;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t c = i / kk;
    int loc = (int)(i % kk);
    int li = loc / nloc, lj = loc % nloc;
    uint64_t r = (uint64_t)cd[c * nloc + li], q = (uint64_t)cd[c * nloc + lj];
    keys[i] = (r << B) | q;
    vals[i] = (uint32_t)i;
  }
}

__global__ void k_face_pair_keys(int64_t nbf, int nloc, int nfd, int B, const int32_t* bf_cell,
                                 const int32_t* bf_local, const int32_t* cd,
                                 const int32_t* face_dofs, uint64_t* keys, uint32_t* vals) {
  int64_t kk = (int64_t)nfd * nfd;
  int64_t n = nbf * kk;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t f = i / kk;
    int loc = (int)(i % kk);
    int a = loc / nfd, b = loc % nfd;
    int c = bf_cell[f], lf = bf_local[f];
    uint64_t r = (uint64_t)cd[(int64_t)c * nloc + face_dofs[lf * nfd + a]];
    uint64_t q = (uint64_t)cd[(int64_t)c * nloc + face_dofs[lf * nfd + b]];
    keys[i] = (r << B) | q;
    vals[i] = (uint32_t)i;
  }
}

// Unique keys -> segment starts, compact keys.
__global__ void k_unique(int64_t n, const uint64_t* keys, const int32_t* flags,
                         const int32_t* segid, int32_t* seg, uint64_t* ukeys) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (flags[i]) {
      seg[segid[i] - 1] = (int32_t)i;
      ukeys[segid[i] - 1] = keys[i];
    }
}

// Full pattern: every row has at least its diagonal, so the first unique key of
// each row gives row_ptr directly.
__global__ void k_rowptr_full(int64_t nnz, const uint64_t* ukeys, int B, int32_t* row_ptr,
                              int32_t* col) {
  uint64_t m = (uint64_t(1) << B) - 1;
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nnz;
       u += (int64_t)gridDim.x * blockDim.x) {
    uint64_t r = ukeys[u] >> B;
    col[u] = (int32_t)(ukeys[u] & m);
    if (u == 0 || (ukeys[u - 1] >> B) != r) row_ptr[r] = (int32_t)u;
  }
}

__global__ void k_row_count(int64_t nnz, const uint64_t* ukeys, int B, int32_t* cnt,
                            int32_t* col) {
  uint64_t m = (uint64_t(1) << B) - 1;
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nnz;
       u += (int64_t)gridDim.x * blockDim.x) {
    col[u] = (int32_t)(ukeys[u] & m);
    atomicAdd(&cnt[ukeys[u] >> B], 1);
  }
}

__global__ void k_diag(int64_t n, const int32_t* rp, const int32_t* col, int32_t* diag) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    diag[r] = find_col(col, rp[r], rp[r + 1], (int)r);
  }
}

__global__ void k_set_i32(int64_t n, int32_t* p, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_set_u8(int64_t n, uint8_t* p, uint8_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_or_u8(int64_t n, const uint8_t* a, uint8_t* b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] |= a[i];
}

__global__ void k_count_u8(int64_t n, const uint8_t* p, unsigned long long* out) {
  unsigned long long s = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += p[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, s);
}

constexpr int TB = 256;

int64_t count_flags(Context* c, const uint8_t* p, int64_t n) {
  DBuf<unsigned long long> cnt(1);
  unsigned long long z = 0;
  SUCEM_HIP(hipMemcpyAsync(cnt.p, &z, sizeof(z), hipMemcpyHostToDevice, c->stream));
  SUCEM_LAUNCH(k_count_u8, grid_for(n, TB, 4096), TB, 0, c->stream, n, p, cnt.p);
  return (int64_t)read_scalar(c, cnt.p);
}

// Number the entities with NV vertices (2 = edges, 3 = faces) that each cell has NL of.
template <int NV, int NL>
int64_t number_entities(Context* c, Mesh* m, DBuf<int32_t>& cell_ent, DBuf<int32_t>& ent_verts,
                        DBuf<int32_t>* isb_out, DBuf<int32_t>* bcell_out,
                        DBuf<int32_t>* blocal_out) {
  const int64_t n = m->nc * NL;
  SUCEM_CHECK(n < (int64_t(1) << 31), SUCEM_ERR_UNSUPPORTED, "mesh too large for 32-bit entity ids");
  const int B = bits_for(m->nv);
  SUCEM_CHECK(B * NV <= 64, SUCEM_ERR_UNSUPPORTED, "too many vertices for packed entity keys");
  DBuf<uint64_t> k0(n), k1(n);
  DBuf<uint32_t> v0(n), v1(n);
  SUCEM_LAUNCH((k_entity_keys<NV, NL>), grid_for(n, TB), TB, 0, c->stream, m->nc, m->cells.p, B,
               k0.p, v0.p);
  sort_pairs(c, k0.p, k1.p, v0.p, v1.p, n, (unsigned)(B * NV));
  k0.release();
  v0.release();
  DBuf<int32_t> flags(n), segid(n);
  SUCEM_LAUNCH(k_heads, grid_for(n, TB), TB, 0, c->stream, n, k1.p, flags.p);
  inclusive_scan_i32(c, flags.p, segid.p, n);
  const int64_t nent = read_scalar(c, segid.p + (n - 1));
  DBuf<int32_t> mark(n), rank_at(n);
  SUCEM_LAUNCH(k_set_i32, grid_for(n, TB), TB, 0, c->stream, n, mark.p, 0);
  SUCEM_LAUNCH(k_mark_first, grid_for(n, TB), TB, 0, c->stream, n, flags.p, v1.p, mark.p);
  exclusive_scan_i32(c, mark.p, rank_at.p, n);
  mark.release();
  cell_ent.alloc(n);
  ent_verts.alloc(nent * NV);
  DBuf<int32_t> seg_start(nent), ent_of_seg(nent);
  SUCEM_LAUNCH((k_assign<NV>), grid_for(n, TB), TB, 0, c->stream, n, k1.p, flags.p, segid.p, v1.p,
               rank_at.p, B, cell_ent.p, ent_verts.p, seg_start.p, ent_of_seg.p);
  SUCEM_LAUNCH(k_scatter_ids, grid_for(n, TB), TB, 0, c->stream, n, segid.p, v1.p, ent_of_seg.p,
               cell_ent.p);
  if (isb_out) {
    isb_out->alloc(nent);
    bcell_out->alloc(nent);
    blocal_out->alloc(nent);
    DBuf<int32_t> bad(1);
    SUCEM_LAUNCH(k_set_i32, 1, 64, 0, c->stream, (int64_t)1, bad.p, 0);
    SUCEM_LAUNCH(k_face_boundary, grid_for(nent, TB), TB, 0, c->stream, nent, n, seg_start.p,
                 ent_of_seg.p, v1.p, isb_out->p, bcell_out->p, blocal_out->p, bad.p);
    SUCEM_CHECK(read_scalar(c, bad.p) == 0, SUCEM_ERR_INVALID,
                "invalid mesh: a face is shared by more than two cells");
  }
  c->sync();
  return nent;
}

// Sort (row, col) pair keys and build a CSR pattern plus the gather map.
struct PatternOut {
  int64_t nnz = 0;
};

int64_t build_pattern(Context* c, int64_t n, DBuf<uint64_t>& keys, DBuf<uint32_t>& vals, int B,
                      int64_t nrows, bool full, DBuf<int32_t>& row_ptr, DBuf<int32_t>& col,
                      DBuf<int32_t>& seg, DBuf<uint32_t>& perm) {
  DBuf<uint64_t> ks(n);
  perm.alloc(n);
  if (n) sort_pairs(c, keys.p, ks.p, vals.p, perm.p, n, (unsigned)(2 * B));
  keys.release();
  vals.release();
  row_ptr.alloc(nrows + 1);
  if (n == 0) {
    SUCEM_LAUNCH(k_set_i32, grid_for(nrows + 1, TB), TB, 0, c->stream, nrows + 1, row_ptr.p, 0);
    seg.alloc(1);
    SUCEM_LAUNCH(k_set_i32, 1, 64, 0, c->stream, (int64_t)1, seg.p, 0);
    col.alloc(0);
    return 0;
  }
  DBuf<int32_t> flags(n), segid(n);
  SUCEM_LAUNCH(k_heads, grid_for(n, TB), TB, 0, c->stream, n, ks.p, flags.p);
  inclusive_scan_i32(c, flags.p, segid.p, n);
  const int64_t nnz = read_scalar(c, segid.p + (n - 1));
  seg.alloc(nnz + 1);
  DBuf<uint64_t> uk(nnz);
  SUCEM_LAUNCH(k_unique, grid_for(n, TB), TB, 0, c->stream, n, ks.p, flags.p, segid.p, seg.p,
               uk.p);
  int32_t n32 = (int32_t)n;
  SUCEM_HIP(hipMemcpyAsync(seg.p + nnz, &n32, sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  col.alloc(nnz);
  if (full) {
    SUCEM_LAUNCH(k_rowptr_full, grid_for(nnz, TB), TB, 0, c->stream, nnz, uk.p, B, row_ptr.p,
                 col.p);
    int32_t z = (int32_t)nnz;
    SUCEM_HIP(hipMemcpyAsync(row_ptr.p + nrows, &z, sizeof(int32_t), hipMemcpyHostToDevice,
                             c->stream));
  } else {
    DBuf<int32_t> cnt(nrows + 1);
    SUCEM_LAUNCH(k_set_i32, grid_for(nrows + 1, TB), TB, 0, c->stream, nrows + 1, cnt.p, 0);
    SUCEM_LAUNCH(k_row_count, grid_for(nnz, TB), TB, 0, c->stream, nnz, uk.p, B, cnt.p, col.p);
    exclusive_scan_i32(c, cnt.p, row_ptr.p, nrows + 1);
  }
  c->sync();
  return nnz;
}

}  // namespace

int64_t build_csr_from_pairs(Context* c, int64_t n, DBuf<uint64_t>& keys, DBuf<uint32_t>& vals,
                             int B, int64_t nrows, DBuf<int32_t>& row_ptr, DBuf<int32_t>& col,
                             DBuf<int32_t>& seg, DBuf<uint32_t>& perm) {
  return build_pattern(c, n, keys, vals, B, nrows, false, row_ptr, col, seg, perm);
}

int key_bits(int64_t n) { return bits_for(n); }

void mesh_create_box(Context* c, Mesh* m, int64_t nx, int64_t ny, int64_t nz,
                     const double scale[3], const double shift[3]) {
  SUCEM_CHECK(nx > 0 && ny > 0 && nz > 0, SUCEM_ERR_INVALID, "box divisions must be positive");
  m->ctx = c;
  m->nv = (nx + 1) * (ny + 1) * (nz + 1);
  m->nc = 6 * nx * ny * nz;
  SUCEM_CHECK(m->nv < (int64_t(1) << 31) && m->nc * 20 < (int64_t(1) << 31),
              SUCEM_ERR_UNSUPPORTED, "box mesh too large");
  m->xyz.alloc(m->nv * 3);
  m->cells.alloc(m->nc * 4);
  SUCEM_LAUNCH(k_box_vertices, grid_for(m->nv, TB), TB, 0, c->stream, nx, ny, nz, scale[0],
               scale[1], scale[2], shift[0], shift[1], shift[2], m->xyz.p);
  SUCEM_LAUNCH(k_box_cells, grid_for(nx * ny * nz, TB), TB, 0, c->stream, nx, ny, nz, m->cells.p);
}

void mesh_translate(Context* c, Mesh* m, const double shift[3]) {
  SUCEM_LAUNCH(k_translate, grid_for(m->nv, TB), TB, 0, c->stream, m->nv, m->xyz.p, shift[0],
               shift[1], shift[2]);
}

void mesh_sort_cells(Context* c, Mesh* m) {
  SUCEM_LAUNCH(k_sort_cells, grid_for(m->nc, TB), TB, 0, c->stream, m->nc, m->cells.p);
}

int64_t mesh_set_materials(Context* c, Mesh* m, const int32_t* region_d, int64_t nreg,
                           const int64_t* ids_h, const double* eps_h, const double* mu_h) {
  SUCEM_CHECK(nreg > 0, SUCEM_ERR_INVALID, "empty material table");
  for (int64_t i = 0; i < nreg; ++i) {
    SUCEM_CHECK(mu_h[i] != 0.0, SUCEM_ERR_INVALID, "mu_r must be nonzero");
    SUCEM_CHECK(std::isfinite(mu_h[i]) && std::isfinite(eps_h[i]), SUCEM_ERR_INVALID,
                "eps_r and mu_r must be finite");
  }
  DBuf<int64_t> ids(nreg);
  DBuf<double> eps(nreg), mu(nreg);
  h2d(c, ids.p, ids_h, nreg);
  h2d(c, eps.p, eps_h, nreg);
  h2d(c, mu.p, mu_h, nreg);
  m->eps_r.alloc(m->nc);
  m->inv_mu_r.alloc(m->nc);
  DBuf<unsigned long long> bad(1);
  unsigned long long none = ~0ull;
  SUCEM_HIP(hipMemcpyAsync(bad.p, &none, sizeof(none), hipMemcpyHostToDevice, c->stream));
  SUCEM_LAUNCH(k_materials, grid_for(m->nc, TB), TB, 0, c->stream, m->nc, region_d, (int)nreg,
               ids.p, eps.p, mu.p, m->eps_r.p, m->inv_mu_r.p, bad.p);
  unsigned long long b = read_scalar(c, bad.p);
  return b == ~0ull ? -1 : (int64_t)b;
}

void space_build(Context* c, Space* s) {
  Mesh* m = s->mesh;
  SUCEM_CHECK(s->order == 1 || s->order == 2, SUCEM_ERR_UNSUPPORTED,
              "only Nedelec orders 1 and 2 are supported");
  const RefData& R = ref_data(s->order);
  s->nloc = R.nloc;
  s->nfd = R.nfd;
  // reference tensors
  s->ref_mass.alloc(R.mass.size());
  s->ref_stiff.alloc(R.stiff.size());
  s->ref_face.alloc(R.face.size());
  s->ref_basis.alloc(R.basis.size() * 2);
  s->ref_face_dofs.alloc(R.face_dofs.size());
  h2d(c, s->ref_mass.p, R.mass.data(), R.mass.size());
  h2d(c, s->ref_stiff.p, R.stiff.data(), R.stiff.size());
  h2d(c, s->ref_face.p, R.face.data(), R.face.size());
  h2d(c, s->ref_basis.p, R.basis.data(), R.basis.size());
  h2d(c, s->ref_basis.p + R.basis.size(), R.curl.data(), R.curl.size());
  h2d(c, s->ref_face_dofs.p, R.face_dofs.data(), R.face_dofs.size());

  // 1. entities
  s->nE = number_entities<2, 6>(c, m, s->cell_edges, s->edges, nullptr, nullptr, nullptr);
  DBuf<int32_t> isb, bcell, blocal;
  s->nF = number_entities<3, 4>(c, m, s->cell_faces, s->faces, &isb, &bcell, &blocal);
  // boundary faces in face-id order
  {
    DBuf<int32_t> pos(s->nF + 1);
    exclusive_scan_i32(c, isb.p, pos.p, s->nF);
    int32_t last_pos = read_scalar(c, pos.p + (s->nF - 1));
    int32_t last_isb = read_scalar(c, isb.p + (s->nF - 1));
    s->nbf = (int64_t)last_pos + last_isb;
    s->bf_cell.alloc(s->nbf);
    s->bf_local.alloc(s->nbf);
    SUCEM_LAUNCH(k_compact_bfaces, grid_for(s->nF, TB), TB, 0, c->stream, s->nF, isb.p, pos.p,
                 bcell.p, blocal.p, s->bf_cell.p, s->bf_local.p);
  }
  // 2. DOF map
  s->ndofs = s->order == 1 ? s->nE : 2 * s->nE + 2 * s->nF;
  SUCEM_CHECK(s->ndofs < (int64_t(1) << 31), SUCEM_ERR_UNSUPPORTED, "too many DOFs");
  s->cell_dofs.alloc(m->nc * s->nloc);
  SUCEM_LAUNCH(k_dofmap, grid_for(m->nc, TB), TB, 0, c->stream, m->nc, s->order, s->nE,
               s->cell_edges.p, s->cell_faces.p, s->cell_dofs.p);
  // 3. PEC mask: DOFs of the closure of every boundary face
  s->pec.alloc(s->ndofs);
  SUCEM_LAUNCH(k_set_u8, grid_for(s->ndofs, TB), TB, 0, c->stream, s->ndofs, s->pec.p, (uint8_t)0);
  if (s->nbf)
    SUCEM_LAUNCH(k_pec, grid_for(s->nbf * s->nfd, TB), TB, 0, c->stream, s->nbf, s->nloc, s->nfd,
                 s->bf_cell.p, s->bf_local.p, s->cell_dofs.p, s->ref_face_dofs.p, s->pec.p);
  s->bdof.alloc(s->ndofs);
  SUCEM_HIP(hipMemcpyAsync(s->bdof.p, s->pec.p, s->ndofs, hipMemcpyDeviceToDevice, c->stream));
  s->npec = count_flags(c, s->pec.p, s->ndofs);
  // 4. CSR pattern of all cell-local pairs + the assembly gather map
  const int B = bits_for(s->ndofs);
  {
    const int64_t n = m->nc * s->nloc * s->nloc;
    SUCEM_CHECK(n < (int64_t(1) << 32), SUCEM_ERR_UNSUPPORTED, "too many element contributions");
    SUCEM_CHECK(n < (int64_t(1) << 31), SUCEM_ERR_UNSUPPORTED, "pattern exceeds 32-bit offsets");
    DBuf<uint64_t> keys(n);
    DBuf<uint32_t> vals(n);
    SUCEM_LAUNCH(k_cell_pair_keys, grid_for(n, TB), TB, 0, c->stream, m->nc, s->nloc, B,
                 s->cell_dofs.p, keys.p, vals.p);
    s->nnz = build_pattern(c, n, keys, vals, B, s->ndofs, true, s->row_ptr, s->col, s->seg,
                           s->perm);
    s->diag.alloc(s->ndofs);
    SUCEM_LAUNCH(k_diag, grid_for(s->ndofs, TB), TB, 0, c->stream, s->ndofs, s->row_ptr.p,
                 s->col.p, s->diag.p);
  }
  // 5. compact ABC pattern over boundary-face DOF pairs
  {
    const int64_t n = s->nbf * s->nfd * s->nfd;
    DBuf<uint64_t> keys(n);
    DBuf<uint32_t> vals(n);
    if (n)
      SUCEM_LAUNCH(k_face_pair_keys, grid_for(n, TB), TB, 0, c->stream, s->nbf, s->nloc, s->nfd,
                   B, s->bf_cell.p, s->bf_local.p, s->cell_dofs.p, s->ref_face_dofs.p, keys.p,
                   vals.p);
    s->nnz0 = build_pattern(c, n, keys, vals, B, s->ndofs, false, s->s0_ptr, s->s0_col, s->s0_seg,
                            s->s0_perm);
    s->s0_diag.alloc(s->ndofs);
    SUCEM_LAUNCH(k_diag, grid_for(s->ndofs, TB), TB, 0, c->stream, s->ndofs, s->s0_ptr.p,
                 s->s0_col.p, s->s0_diag.p);
  }
  c->sync();
}

void space_mark_pec(Context* c, Space* s, const uint8_t* facet_flag_h, bool union_mode) {
  if (!facet_flag_h) {
    if (union_mode) {
      SUCEM_LAUNCH(k_or_u8, grid_for(s->ndofs, TB, 4096), TB, 0, c->stream, s->ndofs, s->bdof.p,
                   s->pec.p);
    } else {
      SUCEM_HIP(hipMemcpyAsync(s->pec.p, s->bdof.p, s->ndofs, hipMemcpyDeviceToDevice, c->stream));
    }
  } else {
    DBuf<uint8_t> flag(s->nF);
    h2d(c, flag.p, facet_flag_h, s->nF);
    if (!union_mode)
      SUCEM_LAUNCH(k_set_u8, grid_for(s->ndofs, TB), TB, 0, c->stream, s->ndofs, s->pec.p,
                   (uint8_t)0);
    const int64_t nc = s->mesh->nc;
    SUCEM_LAUNCH(k_pec_facets, grid_for(nc * 4, TB), TB, 0, c->stream, nc, s->nloc, s->nfd,
                 s->cell_faces.p, flag.p, s->cell_dofs.p, s->ref_face_dofs.p, s->pec.p);
    c->sync();   // flag is released on return
  }
  s->npec = count_flags(c, s->pec.p, s->ndofs);
  s->pec_version += 1;
}

}  // namespace sucem
