// extern "C" entry points of libsucem_hip (declared in include/sucem_hip.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"
#include "ops.hpp"

namespace sucem {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

double* Context::host_scratch(size_t count) {
  if (count > pinned_n) {
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
    size_t want = std::max<size_t>(count, 4096);
    SUCEM_HIP(hipHostMalloc((void**)&pinned, want * sizeof(double), hipHostMallocDefault));
    pinned_n = want;
  }
  return pinned;
}

// export (AssemblyExport) helpers defined in assemble.hip / mesh_space.hip
void mesh_translate(Context* c, Mesh* m, const double shift[3]);
void export_lhs(Context* c, const System* sys, double k0, double2* out_d);

}  // namespace sucem

using namespace sucem;

#define API_BEGIN try {
#define API_END                                        \
  return SUCEM_OK;                                     \
  }                                                    \
  catch (const sucem::Error& e) {                      \
    set_last_error(e.what());                          \
    return e.code;                                     \
  }                                                    \
  catch (const std::bad_alloc&) {                      \
    set_last_error("host allocation failed");          \
    return SUCEM_ERR_OOM;                              \
  }                                                    \
  catch (const std::exception& e) {                    \
    set_last_error(e.what());                          \
    return SUCEM_ERR_INVALID;                          \
  }

#define NOT_NULL(p) SUCEM_CHECK((p) != nullptr, SUCEM_ERR_INVALID, #p " must not be NULL")

namespace {
template <typename T>
void to_host_i64(Context* c, int64_t* dst, const T* src_d, size_t n) {
  std::vector<T> tmp(n);
  d2h(c, tmp.data(), src_d, n);
  c->sync();
  for (size_t i = 0; i < n; ++i) dst[i] = (int64_t)tmp[i];
}
}  // namespace

extern "C" {

const char* sucem_last_error(void) { return g_last_error.c_str(); }

int sucem_version(void) { return 10000; }

int sucem_ctx_create(int device, void* hip_stream, sucem_ctx** out) {
  API_BEGIN
  NOT_NULL(out);
  int ndev = 0;
  SUCEM_HIP(hipGetDeviceCount(&ndev));
  SUCEM_CHECK(device >= 0 && device < ndev, SUCEM_ERR_INVALID, "no such HIP device");
  auto* h = new sucem_ctx();
  Context& c = h->c;
  c.device = device;
  try {
    SUCEM_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    SUCEM_HIP(hipGetDeviceProperties(&prop, device));
    c.num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    noalloc_init(device);   // device address of the allocation-free staging array
    // Work is issued on exactly the caller's stream (NULL = the device's null
    // stream), so it is ordered with the caller's own copies and kernels.
    c.stream = (hipStream_t)hip_stream;
    c.own_stream = false;
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int sucem_ctx_destroy(sucem_ctx* ctx) {
  API_BEGIN
  if (!ctx) return SUCEM_OK;
  Context& c = ctx->c;
  (void)hipSetDevice(c.device);
  (void)hipStreamSynchronize(c.stream);
  if (c.aux_stream) {
    (void)hipStreamSynchronize(c.aux_stream);
    (void)hipStreamDestroy(c.aux_stream);
  }
  if (c.pinned) (void)hipHostFree(c.pinned);
  if (c.own_stream) (void)hipStreamDestroy(c.stream);
  delete ctx;
  API_END
}

int sucem_ctx_synchronize(sucem_ctx* ctx) {
  API_BEGIN
  NOT_NULL(ctx);
  ctx->c.use();
  ctx->c.sync();
  API_END
}

int sucem_ctx_workspace_bytes(sucem_ctx* ctx, int64_t* bytes_h) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(bytes_h);
  *bytes_h = (int64_t)ctx->c.ws.bytes();
  API_END
}

int sucem_ctx_release_workspace(sucem_ctx* ctx) {
  API_BEGIN
  NOT_NULL(ctx);
  ctx->c.use();
  ctx->c.sync();
  ctx->c.ws.release();
  API_END
}

int sucem_ctx_mem_info(sucem_ctx* ctx, int64_t* free_h, int64_t* total_h) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(free_h); NOT_NULL(total_h);
  ctx->c.use();
  size_t fr = 0, tot = 0;
  SUCEM_HIP(hipMemGetInfo(&fr, &tot));
  *free_h = (int64_t)fr;
  *total_h = (int64_t)tot;
  API_END
}

// ---------------------------------------------------------------- mesh
int sucem_mesh_create(sucem_ctx* ctx, const double* xyz_h, int64_t nv, const int64_t* cells_h,
                      int64_t nc, sucem_mesh** out) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(out); NOT_NULL(xyz_h); NOT_NULL(cells_h);
  SUCEM_CHECK(nv >= 4 && nc >= 1, SUCEM_ERR_INVALID, "mesh needs at least one tetrahedron");
  SUCEM_CHECK(nv < (int64_t(1) << 31) && nc * 20 < (int64_t(1) << 31), SUCEM_ERR_UNSUPPORTED,
              "mesh too large for 32-bit indices");
  Context* c = &ctx->c;
  c->use();
  std::vector<int32_t> cells(nc * 4);
  for (int64_t i = 0; i < nc * 4; ++i) {
    SUCEM_CHECK(cells_h[i] >= 0 && cells_h[i] < nv, SUCEM_ERR_INVALID, "cell vertex out of range");
    cells[i] = (int32_t)cells_h[i];
  }
  auto* h = new sucem_mesh();
  try {
    Mesh& m = h->m;
    m.ctx = c;
    m.nv = nv;
    m.nc = nc;
    m.xyz.alloc(nv * 3);
    m.cells.alloc(nc * 4);
    h2d(c, m.xyz.p, xyz_h, nv * 3);
    h2d(c, m.cells.p, cells.data(), nc * 4);
    mesh_sort_cells(c, &m);
    c->sync();
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int sucem_mesh_create_box(sucem_ctx* ctx, int64_t nx, int64_t ny, int64_t nz,
                          const double* scale3_h, const double* shift3_h, sucem_mesh** out) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(out);
  Context* c = &ctx->c;
  c->use();
  double sc[3] = {1, 1, 1}, sh[3] = {0, 0, 0};
  if (scale3_h) std::memcpy(sc, scale3_h, sizeof(sc));
  if (shift3_h) std::memcpy(sh, shift3_h, sizeof(sh));
  auto* h = new sucem_mesh();
  try {
    mesh_create_box(c, &h->m, nx, ny, nz, sc, sh);
    c->sync();
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int sucem_mesh_info(const sucem_mesh* m, int64_t* nv, int64_t* nc) {
  API_BEGIN
  NOT_NULL(m);
  if (nv) *nv = m->m.nv;
  if (nc) *nc = m->m.nc;
  API_END
}

int sucem_mesh_export(sucem_mesh* m, double* xyz_h, int64_t* cells_h) {
  API_BEGIN
  NOT_NULL(m);
  Context* c = m->m.ctx;
  c->use();
  if (xyz_h) d2h(c, xyz_h, m->m.xyz.p, m->m.nv * 3);
  if (cells_h) to_host_i64(c, cells_h, m->m.cells.p, m->m.nc * 4);
  c->sync();
  API_END
}

int sucem_mesh_translate(sucem_mesh* m, const double* shift3_h) {
  API_BEGIN
  NOT_NULL(m); NOT_NULL(shift3_h);
  Context* c = m->m.ctx;
  c->use();
  mesh_translate(c, &m->m, shift3_h);
  c->sync();
  API_END
}

int sucem_mesh_set_materials(sucem_mesh* m, const int64_t* region_h, int64_t nregions,
                             const int64_t* region_ids_h, const double* eps_r_h,
                             const double* mu_r_h) {
  API_BEGIN
  NOT_NULL(m); NOT_NULL(region_ids_h); NOT_NULL(eps_r_h); NOT_NULL(mu_r_h);
  Context* c = m->m.ctx;
  c->use();
  DBuf<int32_t> reg;
  if (region_h) {
    std::vector<int32_t> r(m->m.nc);
    for (int64_t i = 0; i < m->m.nc; ++i) {
      SUCEM_CHECK(region_h[i] >= INT32_MIN && region_h[i] <= INT32_MAX, SUCEM_ERR_MATERIAL,
                  "Material number " + std::to_string(region_h[i]) + " not found");
      r[i] = (int32_t)region_h[i];
    }
    reg.alloc(m->m.nc);
    h2d(c, reg.p, r.data(), m->m.nc);
  }
  int64_t bad = mesh_set_materials(c, &m->m, reg.p, nregions, region_ids_h, eps_r_h, mu_r_h);
  if (bad >= 0) {
    int64_t rid = region_h ? region_h[bad] : 0;
    throw Error(SUCEM_ERR_MATERIAL, "Material number " + std::to_string(rid) + " not found");
  }
  API_END
}

int sucem_mesh_set_cell_materials(sucem_mesh* m, const double* eps_r_h,
                                  const double* inv_mu_r_h) {
  API_BEGIN
  NOT_NULL(m); NOT_NULL(eps_r_h); NOT_NULL(inv_mu_r_h);
  Context* c = m->m.ctx;
  c->use();
  const int64_t nc = m->m.nc;
  for (int64_t i = 0; i < nc; ++i)
    SUCEM_CHECK(std::isfinite(eps_r_h[i]) && std::isfinite(inv_mu_r_h[i]), SUCEM_ERR_INVALID,
                "cell materials must be finite");
  m->m.eps_r.alloc(nc);
  m->m.inv_mu_r.alloc(nc);
  h2d(c, m->m.eps_r.p, eps_r_h, nc);
  h2d(c, m->m.inv_mu_r.p, inv_mu_r_h, nc);
  c->sync();
  API_END
}

int sucem_mesh_destroy(sucem_mesh* m) {
  API_BEGIN
  if (!m) return SUCEM_OK;
  (void)hipSetDevice(m->m.ctx->device);
  (void)hipStreamSynchronize(m->m.ctx->stream);
  delete m;
  API_END
}

// ---------------------------------------------------------------- space
int sucem_space_create(sucem_ctx* ctx, sucem_mesh* m, int order, sucem_space** out) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(m); NOT_NULL(out);
  Context* c = &ctx->c;
  c->use();
  auto* h = new sucem_space();
  try {
    h->s.ctx = c;
    h->s.mesh = &m->m;
    h->s.order = order;
    space_build(c, &h->s);
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int sucem_space_info(const sucem_space* s, int64_t* info) {
  API_BEGIN
  NOT_NULL(s); NOT_NULL(info);
  const Space& p = s->s;
  info[0] = p.ndofs; info[1] = p.nnz; info[2] = p.nE; info[3] = p.nF;
  info[4] = p.nbf; info[5] = p.nloc; info[6] = p.nnz0; info[7] = p.npec;
  API_END
}

int sucem_space_export_pattern(sucem_space* s, int64_t* indptr_h, int32_t* indices_h) {
  API_BEGIN
  NOT_NULL(s);
  Context* c = s->s.ctx;
  c->use();
  if (indptr_h) to_host_i64(c, indptr_h, s->s.row_ptr.p, s->s.ndofs + 1);
  if (indices_h) d2h(c, indices_h, s->s.col.p, s->s.nnz);
  c->sync();
  API_END
}

int sucem_space_export_dofmap(sucem_space* s, int64_t* cell_dofs_h) {
  API_BEGIN
  NOT_NULL(s); NOT_NULL(cell_dofs_h);
  Context* c = s->s.ctx;
  c->use();
  to_host_i64(c, cell_dofs_h, s->s.cell_dofs.p, s->s.mesh->nc * s->s.nloc);
  API_END
}

int sucem_space_export_entities(sucem_space* s, int64_t* edges_h, int64_t* faces_h) {
  API_BEGIN
  NOT_NULL(s);
  Context* c = s->s.ctx;
  c->use();
  if (edges_h) to_host_i64(c, edges_h, s->s.edges.p, s->s.nE * 2);
  if (faces_h) to_host_i64(c, faces_h, s->s.faces.p, s->s.nF * 3);
  API_END
}

int sucem_space_export_pec(sucem_space* s, uint8_t* mask_h) {
  API_BEGIN
  NOT_NULL(s); NOT_NULL(mask_h);
  Context* c = s->s.ctx;
  c->use();
  d2h(c, mask_h, s->s.pec.p, s->s.ndofs);
  c->sync();
  API_END
}

int sucem_space_mark_pec(sucem_ctx* ctx, sucem_space* s, const uint8_t* facet_flag_h,
                         int union_mode) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(s);
  ctx->c.use();
  space_mark_pec(&ctx->c, &s->s, facet_flag_h, union_mode != 0);
  API_END
}

int sucem_space_destroy(sucem_space* s) {
  API_BEGIN
  if (!s) return SUCEM_OK;
  (void)hipSetDevice(s->s.ctx->device);
  (void)hipStreamSynchronize(s->s.ctx->stream);
  delete s;
  API_END
}

// ---------------------------------------------------------------- assembly
int sucem_assemble(sucem_ctx* ctx, sucem_space* s, unsigned which, sucem_system** out) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(s); NOT_NULL(out);
  SUCEM_CHECK(which != 0 && (which & ~7u) == 0, SUCEM_ERR_INVALID, "bad matrix selection");
  Context* c = &ctx->c;
  c->use();
  auto* h = new sucem_system();
  try {
    assemble_system(c, &s->s, &h->s, which);
    // default operator: matrix-free wherever it applies (5-8x faster than the tuned
    // assembled CSR on a 2.5M-DOF order-2 system), in its patch form (patches of 256
    // cells for FP32 and 192 for FP64 at order 2, 256 at order 1) where the device holds
    // the patch's LDS, else per cell; sucem_system_tune / _set_spmv override
    const unsigned st = SUCEM_MAT_S | SUCEM_MAT_T;
    if ((which & st) == st && (s->s.nloc == 6 || s->s.nloc == 20)) {
      mf_prepare(c, &h->s);
      SpmvCfg f;
      f.mf = 1; f.L = 0; f.U = 1; f.G = op_grid(c, s->s.ndofs, 8);
      h->s.cfg[0] = f;
      h->s.cfg[1] = f;
      const int cb32 = 256, cb64 = s->s.order == 2 ? 192 : 256;
      if (pt_prepare(c, &h->s, cb32) && pt_fits<float>(&s->s, cb32))
        h->s.cfg[0] = pt_cfg(&s->s, cb32, op_grid(c, s->s.pt[2].npd, 16));
      if (pt_prepare(c, &h->s, cb64) && pt_fits<double>(&s->s, cb64))
        h->s.cfg[1] = pt_cfg(&s->s, cb64, op_grid(c, s->s.pt[cb64 == 192 ? 1 : 2].npd, 16));
    }
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int sucem_system_export(sucem_system* sys, unsigned which_one, double* values_h) {
  API_BEGIN
  NOT_NULL(sys); NOT_NULL(values_h);
  Context* c = sys->s.ctx;
  c->use();
  DBuf<double> tmp(sys->s.space->nnz);
  export_values(c, &sys->s, which_one, tmp.p);
  d2h(c, values_h, tmp.p, tmp.n);
  c->sync();
  API_END
}

int sucem_system_export_lhs(sucem_system* sys, double k0, double* values_h) {
  API_BEGIN
  NOT_NULL(sys); NOT_NULL(values_h);
  Context* c = sys->s.ctx;
  c->use();
  DBuf<double2> tmp(sys->s.space->nnz);
  export_lhs(c, &sys->s, k0, tmp.p);
  d2h(c, (double2*)values_h, tmp.p, tmp.n);
  c->sync();
  API_END
}

int sucem_system_destroy(sucem_system* sys) {
  API_BEGIN
  if (!sys) return SUCEM_OK;
  (void)hipSetDevice(sys->s.ctx->device);
  (void)hipStreamSynchronize(sys->s.ctx->stream);
  delete sys;
  API_END
}

int sucem_system_apply(sucem_ctx* ctx, sucem_system* sys, double k0, int precision,
                       int eliminate_pec, int64_t nvec, const void* x_d, void* y_d) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(sys); NOT_NULL(x_d); NOT_NULL(y_d);
  SUCEM_CHECK(nvec >= 0, SUCEM_ERR_INVALID, "nvec must be >= 0");
  Context* c = &ctx->c;
  c->use();
  Coef k = driven_coef(k0);
  if (precision == 64) system_apply<double>(c, &sys->s, k, eliminate_pec != 0, nvec, x_d, y_d);
  else if (precision == 32) system_apply<float>(c, &sys->s, k, eliminate_pec != 0, nvec, x_d, y_d);
  else throw Error(SUCEM_ERR_INVALID, "precision must be 32 or 64");
  API_END
}

int sucem_system_tune(sucem_ctx* ctx, sucem_system* sys, int precision, int reps, int* cfg4_h,
                      double* ms_h) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(sys);
  SUCEM_CHECK(reps > 0, SUCEM_ERR_INVALID, "reps must be positive");
  Context* c = &ctx->c;
  c->use();
  SpmvCfg f;
  if (precision == 32) f = tune_sys<float>(c, &sys->s, reps);
  else if (precision == 64) f = tune_sys<double>(c, &sys->s, reps);
  else throw Error(SUCEM_ERR_INVALID, "precision must be 32 or 64");
  if (cfg4_h) { cfg4_h[0] = f.mf; cfg4_h[1] = f.L; cfg4_h[2] = f.U; cfg4_h[3] = f.G; }
  if (ms_h) *ms_h = f.ms;
  API_END
}

int sucem_system_set_spmv(sucem_ctx* ctx, sucem_system* sys, int precision, const int* cfg4_h) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(sys); NOT_NULL(cfg4_h);
  SUCEM_CHECK(precision == 32 || precision == 64, SUCEM_ERR_INVALID, "precision must be 32 or 64");
  SpmvCfg f;
  f.mf = cfg4_h[0]; f.L = cfg4_h[1]; f.U = cfg4_h[2]; f.G = cfg4_h[3];
  SUCEM_CHECK(f.G > 0 && f.G % 8 == 0 && f.G <= (1 << 20), SUCEM_ERR_INVALID,
              "blocks must be a positive multiple of 8");
  Context* c = &ctx->c;
  c->use();
  if (f.mf == 2) {
    SUCEM_CHECK(f.U == 128 || f.U == 192 || f.U == 256, SUCEM_ERR_INVALID,
                "patch operator: loads (cells per patch) must be 128, 192 or 256");
    SUCEM_CHECK(sys->s.space->order == 2 || f.U == 256, SUCEM_ERR_UNSUPPORTED,
                "patch operator: order-1 patches hold 256 cells");
    SUCEM_CHECK(pt_prepare(c, &sys->s, f.U), SUCEM_ERR_UNSUPPORTED,
                "patch operator: orders 1 and 2, patches of at most 65535 products");
    SUCEM_CHECK(precision == 32 ? pt_fits<float>(sys->s.space, f.U)
                                : pt_fits<double>(sys->s.space, f.U),
                SUCEM_ERR_UNSUPPORTED,
                "the patch operator needs more LDS per workgroup than the device has");
    f = pt_cfg(sys->s.space, f.U, f.G);
  } else if (f.mf == 1) {
    SUCEM_CHECK(sys->s.space->nloc == 6 || sys->s.space->nloc == 20, SUCEM_ERR_UNSUPPORTED,
                "matrix-free operator: orders 1 and 2 only");
    SUCEM_CHECK(f.L == 0 || f.L == 1 || f.L == 4, SUCEM_ERR_INVALID,
                "matrix-free lanes per row must be 0 (LDS-staged rows), 1 or 4");
    SUCEM_CHECK(f.U == 1 || f.U == 2 || f.U == 3, SUCEM_ERR_INVALID,
                "matrix-free loads must be 1 (whole element), 2 (half the rows at a time) or "
                "3 (order-2 element products on the matrix cores, FP32)");
    SUCEM_CHECK(f.U != 3 || (sys->s.space->order == 2 && precision == 32),
                SUCEM_ERR_UNSUPPORTED, "the MFMA element products are order 2, FP32 only");
    mf_prepare(c, &sys->s);
  } else {
    SUCEM_CHECK(f.mf == 0, SUCEM_ERR_INVALID, "matrix_free must be 0, 1 or 2");
    SUCEM_CHECK(f.L == 4 || f.L == 8 || f.L == 16 || f.L == 32, SUCEM_ERR_INVALID,
                "lanes must be 4, 8, 16 or 32");
    SUCEM_CHECK(f.U == 1 || f.U == 2 || f.U == 4, SUCEM_ERR_INVALID, "loads must be 1, 2 or 4");
  }
  sys->s.cfg[precision == 32 ? 0 : 1] = f;
  API_END
}

int sucem_system_patch_info(sucem_ctx* ctx, sucem_system* sys, int cb, int64_t* info6_h) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(sys); NOT_NULL(info6_h);
  SUCEM_CHECK(cb == 128 || cb == 192 || cb == 256, SUCEM_ERR_INVALID,
              "cells per patch must be 128, 192 or 256");
  Context* c = &ctx->c;
  c->use();
  SUCEM_CHECK(pt_prepare(c, &sys->s, cb), SUCEM_ERR_UNSUPPORTED,
              "patch operator: orders 1 and 2, patches of at most 65535 products");
  const PatchMaps& m = sys->s.space->pt[cb == 128 ? 0 : (cb == 192 ? 1 : 2)];
  info6_h[0] = m.np; info6_h[1] = m.nu; info6_h[2] = m.npd; info6_h[3] = m.npe;
  info6_h[4] = m.lds; info6_h[5] = m.maxfr;
  API_END
}

int sucem_system_apply_coef(sucem_ctx* ctx, sucem_system* sys, double cs, double ct, double c0,
                            int precision, int eliminate_pec, int64_t nvec, const void* x_d,
                            void* y_d) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(sys); NOT_NULL(x_d); NOT_NULL(y_d);
  SUCEM_CHECK(nvec >= 0, SUCEM_ERR_INVALID, "nvec must be >= 0");
  Context* c = &ctx->c;
  c->use();
  Coef k;
  k.s = cs; k.t = ct; k.s0 = c0;
  if (precision == 64) system_apply<double>(c, &sys->s, k, eliminate_pec != 0, nvec, x_d, y_d);
  else if (precision == 32) system_apply<float>(c, &sys->s, k, eliminate_pec != 0, nvec, x_d, y_d);
  else throw Error(SUCEM_ERR_INVALID, "precision must be 32 or 64");
  API_END
}

// ---------------------------------------------------------------- sources
int sucem_locate_points(sucem_ctx* ctx, sucem_space* s, int64_t n, const double* pts_h,
                        int64_t* cells_h) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(s); NOT_NULL(pts_h); NOT_NULL(cells_h);
  Context* c = &ctx->c;
  c->use();
  if (n == 0) return SUCEM_OK;
  DBuf<double> pts(n * 3);
  DBuf<int32_t> cell(n);
  h2d(c, pts.p, pts_h, n * 3);
  locate_points(c, &s->s, n, pts.p, cell.p);
  std::vector<int32_t> h(n);
  d2h(c, h.data(), cell.p, n);
  c->sync();
  for (int64_t i = 0; i < n; ++i) {
    if (h[i] == 0x7fffffff)
      throw Error(SUCEM_ERR_OUTSIDE, "point (" + std::to_string(pts_h[3 * i]) + ", " +
                                         std::to_string(pts_h[3 * i + 1]) + ", " +
                                         std::to_string(pts_h[3 * i + 2]) + ") is outside the mesh");
    cells_h[i] = h[i];
  }
  API_END
}

int sucem_basis_at_points(sucem_ctx* ctx, sucem_space* s, int64_t n, const double* pts_h,
                          double* vals_h, int64_t* dofs_h) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(s); NOT_NULL(pts_h); NOT_NULL(vals_h);
  Context* c = &ctx->c;
  c->use();
  if (n == 0) return SUCEM_OK;
  const int nloc = s->s.nloc;
  DBuf<double> pts(n * 3), vals(n * nloc * 3);
  DBuf<int32_t> cell(n);
  h2d(c, pts.p, pts_h, n * 3);
  locate_points(c, &s->s, n, pts.p, cell.p);
  std::vector<int32_t> h(n);
  d2h(c, h.data(), cell.p, n);
  c->sync();
  for (int64_t i = 0; i < n; ++i)
    SUCEM_CHECK(h[i] != 0x7fffffff, SUCEM_ERR_OUTSIDE, "point is outside the mesh");
  basis_at_points(c, &s->s, n, pts.p, cell.p, vals.p);
  d2h(c, vals_h, vals.p, vals.n);
  if (dofs_h) {
    std::vector<int32_t> cd(s->s.mesh->nc * 0 + nloc);
    for (int64_t i = 0; i < n; ++i) {
      d2h(c, cd.data(), s->s.cell_dofs.p + (int64_t)h[i] * nloc, nloc);
      c->sync();
      for (int k = 0; k < nloc; ++k) dofs_h[i * nloc + k] = cd[k];
    }
  }
  c->sync();
  API_END
}

int sucem_rhs_point_sources(sucem_ctx* ctx, sucem_space* s, int64_t nsrc, const double* pos_h,
                            const double* val_h, const int64_t* rhs_h, int64_t nrhs,
                            double scale_re, double scale_im, double* b_d) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(s); NOT_NULL(b_d);
  SUCEM_CHECK(nsrc == 0 || (pos_h && val_h), SUCEM_ERR_INVALID, "missing source arrays");
  Context* c = &ctx->c;
  c->use();
  if (nsrc == 0 || nrhs == 0) return SUCEM_OK;
  std::vector<int32_t> rhs(nsrc, 0);
  for (int64_t i = 0; i < nsrc; ++i) {
    int64_t r = rhs_h ? rhs_h[i] : 0;
    SUCEM_CHECK(r >= 0 && r < nrhs, SUCEM_ERR_INVALID, "source rhs index out of range");
    rhs[i] = (int32_t)r;
  }
  DBuf<double> pos(nsrc * 3);
  DBuf<double2> val(nsrc * 3);
  DBuf<int32_t> rhs_d(nsrc), cell(nsrc);
  h2d(c, pos.p, pos_h, nsrc * 3);
  h2d(c, val.p, (const double2*)val_h, nsrc * 3);
  h2d(c, rhs_d.p, rhs.data(), nsrc);
  locate_points(c, &s->s, nsrc, pos.p, cell.p);
  std::vector<int32_t> h(nsrc);
  d2h(c, h.data(), cell.p, nsrc);
  c->sync();
  for (int64_t i = 0; i < nsrc; ++i)
    SUCEM_CHECK(h[i] != 0x7fffffff, SUCEM_ERR_OUTSIDE, "point source is outside the mesh");
  rhs_point_sources(c, &s->s, nsrc, pos.p, val.p, rhs_d.p, cell.p, nrhs,
                    make_double2(scale_re, scale_im), (double2*)b_d);
  c->sync();
  API_END
}

int sucem_rhs_point_sources_noalloc(sucem_ctx* ctx, sucem_space* s, int64_t nsrc,
                                    const double* pos_h, const double* val_h,
                                    const int64_t* rhs_h, int64_t nrhs, double scale_re,
                                    double scale_im, double* b_d) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(s); NOT_NULL(b_d);
  SUCEM_CHECK(nsrc == 0 || (pos_h && val_h), SUCEM_ERR_INVALID, "missing source arrays");
  SUCEM_CHECK(nsrc >= 0 && nsrc <= kNoAllocSources, SUCEM_ERR_INVALID,
              "the allocation-free point-source call takes at most 4096 sources");
  for (int64_t i = 0; rhs_h && i < nsrc; ++i)
    SUCEM_CHECK(rhs_h[i] >= 0 && rhs_h[i] < nrhs, SUCEM_ERR_INVALID,
                "source rhs index out of range");
  Context* c = &ctx->c;
  c->use();
  rhs_point_sources_noalloc(c, &s->s, nsrc, pos_h, val_h, rhs_h, nrhs,
                            make_double2(scale_re, scale_im), (double2*)b_d);
  API_END
}

// ---------------------------------------------------------------- solves
void sucem_solve_opts_default(sucem_solve_opts* o) {
  if (!o) return;
  o->method = SUCEM_METHOD_COCG;
  o->precision = 64;
  o->precond = SUCEM_PRECOND_JACOBI;
  o->max_iter = 20000;
  o->check_every = 8;
  o->eliminate_pec = 0;
  o->tol = 1e-8;
  o->ru_delta = 0.01;
  o->gmres_restart = 20;
  o->paired = 0;
  o->precond_shift = -1.0;
}

// auto_shift: the preconditioner's complex shift used when the caller leaves it negative
// (kDrivenShift for the driven entry points, 0 for general coefficients and generic
// matrices).  kDrivenShift: on the 40^3 order-2 driven cube FP64 COCG takes 2,032 /
// 1,664 / 1,584 / 1,552 / 1,552 / 1,568 / 1,632 / 1,600 / 1,696 iterations at beta = 0 /
// 0.1 / 0.2 / 0.25 / 0.3 / 0.4 / 0.5 / 0.75 / 1; the 40^3 order-1 sweep 796 / 768 / 820
// (mean) at 0 / 0.25 / 0.5.
constexpr double kDrivenShift = 0.25;
static SolveOpts convert(const sucem_solve_opts* o, double auto_shift) {
  SolveOpts s;
  s.precond_shift = auto_shift;
  if (o) {
    if (o->precond_shift >= 0) s.precond_shift = o->precond_shift;
    s.method = o->method;
    s.precision = o->precision;
    s.precond = o->precond;
    s.max_iter = o->max_iter;
    s.check_every = o->check_every;
    s.eliminate_pec = o->eliminate_pec;
    s.tol = o->tol;
    s.ru_delta = o->ru_delta;
    s.restart = o->gmres_restart;
    s.paired = o->paired;
  }
  SUCEM_CHECK(s.tol > 0 && s.ru_delta > 0 && s.ru_delta < 1, SUCEM_ERR_INVALID,
              "tol must be > 0 and ru_delta in (0, 1)");
  SUCEM_CHECK(s.precond_shift >= 0 && s.precond_shift <= 10, SUCEM_ERR_INVALID,
              "precond_shift must be in [0, 10] (or negative for the default)");
  SUCEM_CHECK(s.paired == 0 || s.paired == 1, SUCEM_ERR_INVALID, "paired must be 0 or 1");
  SUCEM_CHECK(s.precond == SUCEM_PRECOND_NONE || s.precond == SUCEM_PRECOND_JACOBI ||
                  s.precond == SUCEM_PRECOND_GRADIENT || s.precond == SUCEM_PRECOND_AUXILIARY,
              SUCEM_ERR_UNSUPPORTED, "unknown preconditioner");
  return s;
}

static void convert_stats(const SolveStats& s, sucem_solve_stats* out) {
  if (!out) return;
  out->status = s.status;
  out->iterations = s.iterations;
  out->reliable_updates = s.reliable_updates;
  out->stagnated = s.stagnated;
  out->rel_residual = s.rel_residual;
  out->elapsed_ms = s.elapsed_ms;
  out->spmv_lo = s.spmv_lo;
  out->spmv_hi = s.spmv_hi;
  out->op_ms = s.op_ms;
  out->op_samples = s.op_samples;
  out->fp64_finish = s.fp32_failed;
  out->batch_width = s.batch_width;
}

int sucem_solve_driven(sucem_ctx* ctx, sucem_system* sys, double k0, const double* b_d,
                       double* x_d, const sucem_solve_opts* opts, sucem_solve_stats* stats,
                       double* res_hist_h, int64_t hist_cap) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(sys); NOT_NULL(b_d); NOT_NULL(x_d);
  Context* c = &ctx->c;
  c->use();
  SolveStats st;
  solve_system(c, &sys->s, driven_coef(k0), (const double2*)b_d, (double2*)x_d, convert(opts, kDrivenShift), st,
               res_hist_h, hist_cap);
  convert_stats(st, stats);
  API_END
}

int sucem_solve_driven_batch(sucem_ctx* ctx, sucem_system* sys, int64_t nb, const double* k0_h,
                             const double* B_d, double* X_d, const sucem_solve_opts* opts,
                             sucem_solve_stats* stats_h) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(sys); NOT_NULL(k0_h); NOT_NULL(B_d); NOT_NULL(X_d);
  SUCEM_CHECK(nb > 0, SUCEM_ERR_INVALID, "nb must be positive");
  Context* c = &ctx->c;
  c->use();
  std::vector<SolveStats> st(nb);
  SolveOpts o = convert(opts, kDrivenShift);
  if (!opts) o.precond = SUCEM_PRECOND_AUXILIARY;
  SUCEM_CHECK(!o.paired, SUCEM_ERR_UNSUPPORTED, "driven systems are complex: no paired solves");
  if (o.method == SUCEM_METHOD_COCG && o.precond == SUCEM_PRECOND_AUXILIARY &&
      sys_supports_sweep(c, &sys->s))
    solve_system_sweep(c, &sys->s, nb, k0_h, (const double2*)B_d, (double2*)X_d, o, 0, true,
                       st.data());
  else
    solve_system_batch(c, &sys->s, nb, k0_h, (const double2*)B_d, (double2*)X_d, o, st.data());
  if (stats_h)
    for (int64_t j = 0; j < nb; ++j) convert_stats(st[j], stats_h + j);
  API_END
}

int sucem_solve_driven_sweep(sucem_ctx* ctx, sucem_system* sys, int64_t nb, const double* k0_h,
                             const double* B_d, double* X_d, const sucem_solve_opts* opts,
                             int width, int use_graph, sucem_solve_stats* stats_h) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(sys); NOT_NULL(k0_h); NOT_NULL(B_d); NOT_NULL(X_d);
  SUCEM_CHECK(nb > 0, SUCEM_ERR_INVALID, "nb must be positive");
  Context* c = &ctx->c;
  c->use();
  SolveOpts o = convert(opts, kDrivenShift);
  if (!opts) o.precond = SUCEM_PRECOND_AUXILIARY;
  SUCEM_CHECK(!o.paired, SUCEM_ERR_UNSUPPORTED, "driven systems are complex: no paired solves");
  std::vector<SolveStats> st(nb);
  solve_system_sweep(c, &sys->s, nb, k0_h, (const double2*)B_d, (double2*)X_d, o, width,
                     use_graph != 0, st.data());
  if (stats_h)
    for (int64_t j = 0; j < nb; ++j) convert_stats(st[j], stats_h + j);
  API_END
}

int sucem_solve_system(sucem_ctx* ctx, sucem_system* sys, double cs, double ct, double c0,
                       const double* b_d, double* x_d, const sucem_solve_opts* opts,
                       sucem_solve_stats* stats, double* res_hist_h, int64_t hist_cap) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(sys); NOT_NULL(b_d); NOT_NULL(x_d);
  Context* c = &ctx->c;
  c->use();
  Coef k;
  k.s = cs; k.t = ct; k.s0 = c0;
  SolveStats st;
  solve_system(c, &sys->s, k, (const double2*)b_d, (double2*)x_d, convert(opts, 0.0), st, res_hist_h,
               hist_cap);
  convert_stats(st, stats);
  API_END
}

int sucem_system_precond_apply(sucem_ctx* ctx, sucem_system* sys, double cs, double ct,
                               double c0, double shift, int precision, int eliminate_pec,
                               const void* r_d, void* z_d) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(sys); NOT_NULL(r_d); NOT_NULL(z_d);
  Context* c = &ctx->c;
  c->use();
  Coef k;
  k.s = cs; k.t = ct; k.s0 = c0;
  precond_apply(c, &sys->s, k, shift, precision, eliminate_pec != 0, r_d, z_d);
  API_END
}

int sucem_surface_flux(sucem_ctx* ctx, sucem_space* V, const double* x_d, double k0, int use_mu,
                       double* flux) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(V); NOT_NULL(x_d); NOT_NULL(flux);
  SUCEM_CHECK(k0 > 0, SUCEM_ERR_INVALID, "k0 must be positive");
  ctx->c.use();
  *flux = surface_flux(&ctx->c, &V->s, x_d, k0, use_mu != 0);
  API_END
}

int sucem_variational_flux(sucem_ctx* ctx, sucem_space* V, const double* x_d, double k0,
                           double* flux) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(V); NOT_NULL(x_d); NOT_NULL(flux);
  SUCEM_CHECK(k0 > 0, SUCEM_ERR_INVALID, "k0 must be positive");
  ctx->c.use();
  variational_flux(&ctx->c, &V->s, x_d, k0, flux);
  API_END
}

int sucem_ntff(sucem_ctx* ctx, sucem_space* V, const double* x_d, double k0, int variational,
               int64_t n, const double* theta_h, const double* phi_h, double* E_ff_h) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(V); NOT_NULL(x_d);
  SUCEM_CHECK(k0 > 0, SUCEM_ERR_INVALID, "k0 must be positive");
  SUCEM_CHECK(n >= 0 && n <= 65535, SUCEM_ERR_INVALID, "direction count must be in [0, 65535]");
  if (n == 0) return SUCEM_OK;
  NOT_NULL(theta_h); NOT_NULL(phi_h); NOT_NULL(E_ff_h);
  ctx->c.use();
  ntff(&ctx->c, &V->s, x_d, k0, variational != 0, n, theta_h, phi_h, E_ff_h);
  API_END
}

void sucem_eig_opts_default(sucem_eig_opts* o) {
  if (!o) return;
  EigOpts d;
  o->ncv = d.ncv;
  o->max_restarts = d.max_restarts;
  o->inner_precision = d.inner_precision;
  o->inner_max_iter = d.inner_max_iter;
  o->tol = d.tol;
  o->inner_tol = d.inner_tol;
  o->seed = d.seed;
  o->block = d.block;
  o->pad_ = 0;
}

int sucem_eigs(sucem_ctx* ctx, sucem_system* sys, double sigma, int nev,
               const sucem_eig_opts* opts, double* k2_h, double* resid_h, double* vecs_d,
               sucem_eig_stats* stats) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(sys); NOT_NULL(k2_h);
  SUCEM_CHECK((sys->s.which & (SUCEM_MAT_S | SUCEM_MAT_T)) == (SUCEM_MAT_S | SUCEM_MAT_T),
              SUCEM_ERR_INVALID, "the eigen solver needs S and T");
  Context* c = &ctx->c;
  c->use();
  EigOpts o;
  if (opts) {
    o.ncv = opts->ncv;
    o.max_restarts = opts->max_restarts;
    o.inner_precision = opts->inner_precision;
    o.inner_max_iter = opts->inner_max_iter;
    o.tol = opts->tol;
    o.inner_tol = opts->inner_tol;
    o.seed = opts->seed;
    o.block = opts->block;
  }
  SUCEM_CHECK(o.block >= 0 && o.block <= 2, SUCEM_ERR_INVALID,
              "Lanczos block size must be 0 (automatic), 1 or 2");
  SUCEM_CHECK(o.inner_precision == 32 || o.inner_precision == 64, SUCEM_ERR_INVALID,
              "inner precision must be 32 or 64");
  SUCEM_CHECK(o.tol > 0 && o.inner_tol > 0 && o.max_restarts >= 0 && o.inner_max_iter > 0,
              SUCEM_ERR_INVALID, "bad eigen options");
  EigStats st;
  eigs_system(c, &sys->s, sigma, nev, o, k2_h, resid_h, (double2*)vecs_d, st);
  if (stats) {
    stats->status = st.status;
    stats->nconv = st.nconv;
    stats->restarts = st.restarts;
    stats->lanczos_steps = st.lanczos_steps;
    stats->inner_failures = st.inner_failures;
    stats->inner_stagnated = st.inner_stagnated;
    stats->inner_fp64_from = st.inner_fp64_from;
    stats->block = st.block;
    stats->inner_iterations = st.inner_iterations;
    stats->max_ritz_residual = st.max_ritz_residual;
    stats->elapsed_ms = st.elapsed_ms;
  }
  API_END
}

int sucem_csr_create(sucem_ctx* ctx, int64_t n, int64_t nnz, const int64_t* indptr_h,
                     const int32_t* indices_h, const double* values_h, sucem_csr** out) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(indptr_h); NOT_NULL(out);
  SUCEM_CHECK(n > 0 && nnz >= 0 && nnz < (int64_t(1) << 31), SUCEM_ERR_INVALID, "bad CSR sizes");
  SUCEM_CHECK(indptr_h[0] == 0 && indptr_h[n] == nnz, SUCEM_ERR_INVALID, "bad CSR indptr");
  Context* c = &ctx->c;
  c->use();
  std::vector<int32_t> rp(n + 1), diag(n, -1);
  for (int64_t i = 0; i <= n; ++i) rp[i] = (int32_t)indptr_h[i];
  for (int64_t i = 0; i < n; ++i) {
    SUCEM_CHECK(indptr_h[i + 1] >= indptr_h[i], SUCEM_ERR_INVALID, "CSR indptr not monotone");
    for (int64_t j = indptr_h[i]; j < indptr_h[i + 1]; ++j) {
      SUCEM_CHECK(indices_h[j] >= 0 && indices_h[j] < n, SUCEM_ERR_INVALID, "CSR column out of range");
      if (indices_h[j] == i && diag[i] < 0) diag[i] = (int32_t)j;
    }
  }
  std::vector<float2> v32(nnz);
  for (int64_t j = 0; j < nnz; ++j)
    v32[j] = make_float2((float)values_h[2 * j], (float)values_h[2 * j + 1]);
  auto* h = new sucem_csr();
  try {
    Csr& A = h->a;
    A.ctx = c;
    A.n = n;
    A.nnz = nnz;
    A.row_ptr.alloc(n + 1);
    A.col.alloc(nnz);
    A.val64.alloc(nnz);
    A.val32.alloc(nnz);
    A.diag.alloc(n);
    h2d(c, A.row_ptr.p, rp.data(), n + 1);
    h2d(c, A.col.p, indices_h, nnz);
    h2d(c, A.val64.p, (const double2*)values_h, nnz);
    h2d(c, A.val32.p, v32.data(), nnz);
    h2d(c, A.diag.p, diag.data(), n);
    c->sync();
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int sucem_csr_apply(sucem_ctx* ctx, sucem_csr* A, const double* x_d, double* y_d) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(A); NOT_NULL(x_d); NOT_NULL(y_d);
  Context* c = &ctx->c;
  c->use();
  csr_apply64(c, &A->a, (const double2*)x_d, (double2*)y_d);
  API_END
}

int sucem_csr_solve(sucem_ctx* ctx, sucem_csr* A, const double* b_d, double* x_d,
                    const sucem_solve_opts* opts, sucem_solve_stats* stats, double* res_hist_h,
                    int64_t hist_cap) {
  API_BEGIN
  NOT_NULL(ctx); NOT_NULL(A); NOT_NULL(b_d); NOT_NULL(x_d);
  Context* c = &ctx->c;
  c->use();
  SolveStats st;
  solve_csr(c, &A->a, (const double2*)b_d, (double2*)x_d, convert(opts, 0.0), st, res_hist_h, hist_cap);
  convert_stats(st, stats);
  API_END
}

int sucem_csr_destroy(sucem_csr* A) {
  API_BEGIN
  if (!A) return SUCEM_OK;
  (void)hipSetDevice(A->a.ctx->device);
  (void)hipStreamSynchronize(A->a.ctx->stream);
  delete A;
  API_END
}

}  // extern "C"
