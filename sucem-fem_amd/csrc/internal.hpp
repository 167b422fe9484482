// Device-resident objects behind the opaque C-ABI handles.
//
// HBM layout (all arrays device-resident, structure-of-arrays):
//   mesh   : xyz double[nv*3], cells int32[nc*4] (vertex ids sorted per cell),
//            eps_r / inv_mu_r double[nc]
//   space  : cell_dofs int32[nc*nloc]; edges int32[nE*2]; faces int32[nF*3];
//            CSR pattern row_ptr int32[N+1], col int32[nnz], diag int32[N];
//            assembly gather map seg int32[nnz+1] + perm uint32[nc*nloc*nloc]
//            (contribution id = cell*nloc^2 + i*nloc + j, sorted by (row, col));
//            boundary faces bf_cell / bf_local int32[nbf]; pec uint8[N];
//            compact ABC pattern s0_ptr int32[N+1] (rows without boundary DOFs
//            are empty), s0_col int32[nnz0], s0_seg/s0_perm like above
//   system : st64 double2[nnz] = (S, T) interleaved, st32 float2[nnz] (same
//            values rounded), s0_64 double[nnz0], s0_32 float[nnz0]
#pragma once
#include <mutex>

#include "common.hpp"

namespace sucem {

struct Mesh {
  Context* ctx = nullptr;
  int64_t nv = 0, nc = 0;
  DBuf<double> xyz;       // nv*3
  DBuf<int32_t> cells;    // nc*4, sorted
  DBuf<double> eps_r;     // nc
  DBuf<double> inv_mu_r;  // nc
};

// Patch operator maps for patches of cb consecutive cells (in Morton order).  Per patch
// its DOFs ("patch DOFs", global ids ascending); per cell slot (cell rank r, local DOF l),
// at lp[l * nc + r], the slot's patch-DOF index (low 16 bits) and its place among the
// patch's products sorted by patch DOF (high 16 bits); per patch DOF the end of its
// product segment and where its sum goes (-1: the DOF lies in this patch only, its sum is
// the result; else its slot in the partial-sum buffer, which is sorted by DOF and then
// patch); per boundary face its products' places; the DOFs shared by patches with the
// segments of their partial sums.
struct PatchMaps {
  int cb = 0;                           // cells per patch; 0: not built
  int64_t np = 0, nu = 0, npd = 0, npe = 0;
  int lds = 0;                          // largest max(patch DOFs, patch products)
  int maxfr = 0;                        // largest number of (face, row) pairs of a patch
  DBuf<uint32_t> lp;                    // nloc * nc
  DBuf<int32_t> uptr, dof, out, fptr, qptr, qdof;
  DBuf<uint16_t> seg, fpos;
  // per DOF: -1 (the DOF lies in one patch) or (first partial sum << 4) | their count,
  // for consumers that add a shared DOF's partial sums themselves (empty when a DOF is
  // shared by more than 15 patches)
  DBuf<int32_t> qcode;
};

struct Space {
  Context* ctx = nullptr;
  std::mutex lazy;                // guards the data built on first use (gradient, matrix-free)
  Mesh* mesh = nullptr;
  int order = 1, nloc = 6, nfd = 3;
  int64_t nE = 0, nF = 0, nbf = 0, ndofs = 0, nnz = 0, nnz0 = 0, npec = 0;
  DBuf<int32_t> cell_dofs;        // nc*nloc
  DBuf<int32_t> cell_edges;       // nc*6
  DBuf<int32_t> cell_faces;       // nc*4
  DBuf<int32_t> edges;            // nE*2
  DBuf<int32_t> faces;            // nF*3
  DBuf<int32_t> bf_cell, bf_local;
  DBuf<uint8_t> pec;              // ndofs: essential (PEC) DOFs, default = bdof
  DBuf<uint8_t> bdof;             // ndofs: DOFs of the closure of the boundary faces
  int pec_version = 0;            // bumped whenever pec changes
  DBuf<int32_t> row_ptr, col, diag;
  DBuf<int32_t> seg;
  DBuf<uint32_t> perm;
  DBuf<int32_t> s0_ptr, s0_col, s0_diag, s0_seg;
  DBuf<uint32_t> s0_perm;
  // reference tensors on device
  DBuf<double> ref_mass, ref_stiff, ref_face, ref_basis;
  DBuf<int32_t> ref_face_dofs;
  // discrete gradient G (N x nlag) and its transpose, built on first use
  int64_t nlag = 0, nnz_g = 0;
  DBuf<int32_t> g_ptr, g_col, gt_ptr, gt_col;
  DBuf<double> g_val, gt_val;
  DBuf<float> g_val32, gt_val32;
  // the same entries packed as (column << 4) | (k & 15): value = k / g_den with a small
  // integer k (the local gradient entries are k/3 for order 2, k for order 1)
  DBuf<uint32_t> g_pk, gt_pk;
  double g_den = 1.0;
  // nodal interpolation Pi of P1 vector fields (N x 3 nv, built on first use): Pi^T stored
  // per (vertex, DOF) pair with its 3 components (SoA: val[d * nnz_pi + entry]) for the
  // set-up of diag(Pi^T A Pi); the preconditioner applies Pi and Pi^T without stored values
  // from these: vertex ->
  // edge entries (e << 1) | [second vertex], vertex -> face entries f << 2, coordinates
  // rounded to FP32 for the FP32 kernels
  int64_t nnz_pi = 0;
  DBuf<int32_t> pit_ptr, pit_col;
  DBuf<double> pit_val;
  DBuf<int32_t> ve_ptr, ve_ent, vf_ptr, vf_ent;
  // the discrete gradient without stored values: face -> edge ids ((p,q), (p,r), (q,r)) and
  // edge -> face entries (f << 2) | position of the edge in the face
  DBuf<int32_t> face_edges, ef_ptr, ef_ent;
  DBuf<float> xyz32;
  // matrix-free operator (built on first use): DOF ids per cell as SoA; the products of
  // cells (cell * nloc + local) and boundary faces (face * nfd + face DOF) sorted by DOF:
  // segment pointers per DOF and each product's position in that order
  DBuf<int32_t> mf_cdT, mf_dptr, mf_posT, mf_posF;
  DBuf<uint16_t> mf_afrag;        // order 2: fp16 operand fragments of the MFMA element kernel
  // patch operator (built on first use): cells in Morton order of their centroids
  // (pt_cell: rank -> cell) and the boundary faces in the order of their cells' ranks
  // (pt_face: original id, pt_fmeta: (cell rank << 2) | local face); per patch size, the
  // maps of PatchMaps
  int pt_state = 0;                     // 0 not built, 1 built, -1 space unsuited
  int pt_max_lds = 0;                   // LDS bytes per workgroup of the device
  DBuf<int32_t> pt_cell, pt_face, pt_fmeta;
  PatchMaps pt[3];                      // patches of 128, 192, 256 cells
};

// SpMV launch configuration: lanes per row, loads per lane, blocks (0 = heuristic).
struct SpmvCfg {
  int L = 0, U = 0, G = 0;
  int mf = 0;      // 1: matrix-free operator (G = grid of its gather kernel),
                   // 2: patch operator (U = cells per patch, G = grid of its
                   //    partial-sum kernel)
  int P = 0;       // block partials an epilogue launch writes (0: G)
  double ms = 0;
  int parts() const { return P > 0 ? P : G; }
};

struct System {
  Context* ctx = nullptr;
  SpmvCfg cfg[2];  // tuned FP32 / FP64 operator
  // diag(G^T S G), diag(G^T T G), diag(G^T S0 G) per Lagrange node, without / with the
  // PEC rows removed (built on first use by the gradient preconditioner)
  DBuf<double> gdiag[2];
  int gdiag_pec_version = -1;     // Space::pec_version gdiag[1] was built with
  // the same for the vector nodal term: diag(Pi^T {S, T, S0} Pi) per (vertex, component)
  DBuf<double> ndiag[2];
  int ndiag_pec_version = -1;
  Space* space = nullptr;
  unsigned which = 0;
  DBuf<double2> st64;
  DBuf<float2> st32;
  DBuf<double> s0_64;
  DBuf<float> s0_32;
  // matrix-free operator: per-cell coefficients (SoA, 12 per cell) and the element buffer
  bool mf_ready = false;
  DBuf<double> mf_coef64, mf_face64;    // 12 per cell / 6 per boundary face, SoA
  DBuf<float> mf_coef32, mf_face32;
  DBuf<double2> mf_yl;
  // patch operator: the coefficients in patch order and the partial-sum buffer
  bool pt_ready = false;
  DBuf<double> pt_coef64, pt_face64;
  DBuf<float> pt_coef32, pt_face32;
  DBuf<double2> pt_part;
};

// Generic complex CSR matrix (user-supplied, e.g. from scipy).
struct Csr {
  Context* ctx = nullptr;
  int64_t n = 0, nnz = 0;
  DBuf<int32_t> row_ptr, col;
  DBuf<double2> val64;
  DBuf<float2> val32;
  DBuf<int32_t> diag;
  SpmvCfg cfg;
};

// mesh_space.hip
void mesh_create_box(Context* c, Mesh* m, int64_t nx, int64_t ny, int64_t nz,
                     const double scale[3], const double shift[3]);
void mesh_sort_cells(Context* c, Mesh* m);
int64_t mesh_set_materials(Context* c, Mesh* m, const int32_t* region_d, int64_t nreg,
                           const int64_t* ids_h, const double* eps_h, const double* mu_h);
void space_build(Context* c, Space* s);

int64_t build_csr_from_pairs(Context* c, int64_t n, DBuf<uint64_t>& keys, DBuf<uint32_t>& vals,
                             int B, int64_t nrows, DBuf<int32_t>& row_ptr, DBuf<int32_t>& col,
                             DBuf<int32_t>& seg, DBuf<uint32_t>& perm);
int key_bits(int64_t n);
// PEC DOFs = closure of the facets with facet_flag_h[f] != 0 (nF flags; null = the
// boundary facets); union_mode ORs into the current mask instead of replacing it.
void space_mark_pec(Context* c, Space* s, const uint8_t* facet_flag_h, bool union_mode);

// assemble.hip
void assemble_system(Context* c, Space* s, System* sys, unsigned which);
void export_values(Context* c, const System* sys, unsigned which_one, double* out_d);

// Coefficients of the operator y = s*S x + t*T x + i*s0*S0 x.  The driven
// operator S - k0^2 T + i k0 S0 is {1, -k0^2, k0}.
struct Coef {
  double s = 1.0, t = 0.0, s0 = 0.0;
  // preconditioner diagonals only: imaginary part of the T coefficient (a complex shift
  // of the operator the diagonals are built from; the operator itself never uses it)
  double ti = 0.0;
};
inline Coef driven_coef(double k0) { Coef k; k.s = 1.0; k.t = -k0 * k0; k.s0 = k0; return k; }

// ops.hip
template <typename R>
void system_apply(Context* c, const System* sys, Coef k, bool eliminate_pec, int64_t nvec,
                  const void* x, void* y);
void csr_apply64(Context* c, const Csr* A, const double2* x, double2* y);
void locate_points(Context* c, const Space* s, int64_t n, const double* pts_d, int32_t* cell_d);
void rhs_point_sources(Context* c, const Space* s, int64_t nsrc, const double* pos_d,
                       const double2* val_d, const int32_t* rhs_d, const int32_t* cell_d,
                       int64_t nrhs, double2 scale, double2* b_d);
void basis_at_points(Context* c, const Space* s, int64_t n, const double* pts_d,
                     const int32_t* cell_d, double* vals_d);
// allocation-free variant of the point-source RHS: staging in a fixed device array
// (noalloc_init once per device with that device current, at context creation)
constexpr int64_t kNoAllocSources = 4096;
void noalloc_init(int device);
void rhs_point_sources_noalloc(Context* c, const Space* s, int64_t nsrc, const double* pos_h,
                               const double* val_h, const int64_t* rhs_h, int64_t nrhs,
                               double2 scale, double2* b_d);

// krylov.hip
struct SolveOpts {
  int method = 0;        // SUCEM_METHOD_*
  int precision = 64;    // 64 or 32 (32 = FP32 inner with FP64 reliable updates)
  int precond = 1;       // SUCEM_PRECOND_*
  int max_iter = 10000;
  int check_every = 8;
  double tol = 1e-8;
  double ru_delta = 0.01;
  int eliminate_pec = 0;
  int warm = 0;          // x holds the initial guess (otherwise x0 = 0)
  int restart = 20;      // GMRES restart length
  // complex shift beta of the preconditioner's operator: its diagonals are built from
  // cs S + ct (1 - i beta) T + i c0 S0 (driven Helmholtz systems converge faster);
  // < 0: not resolved yet (the entry points set 0.5 for driven systems, 0 otherwise)
  double precond_shift = -1.0;
  // b and x hold two real systems, one in the real and one in the imaginary parts, solved
  // as two independent real COCG recursions sharing each operator and preconditioner pass
  // (real coefficients only: c0 = 0, no preconditioner shift; the patch operator; the
  // residuals reported are the larger of the two relative residuals)
  int paired = 0;
};
struct SolveStats {
  int status = 0, iterations = 0, reliable_updates = 0;
  double rel_residual = 0, elapsed_ms = 0;
  int64_t spmv_lo = 0, spmv_hi = 0;
  double op_ms = 0;         // mean device time of a sampled working-precision operator
  int64_t op_samples = 0;
  int stagnated = 0;        // stopped at the floor of the true residual (status > 0)
  int fp32_failed = 0;      // FP32 cycles stopped making progress; finished in FP64
  int batch_width = 1;      // systems solved together with this one (batched sweep)
};
void solve_system(Context* c, const System* sys, Coef k, const double2* b, double2* x,
                  const SolveOpts& o, SolveStats& st, double* hist_h, int64_t hist_cap);
void precond_apply(Context* c, const System* sys, Coef k, double shift, int precision,
                   bool eliminate_pec, const void* r, void* z);
void solve_csr(Context* c, const Csr* A, const double2* b, double2* x, const SolveOpts& o,
               SolveStats& st, double* hist_h, int64_t hist_cap);

// hiptmair.hip: additive gradient-space preconditioner
//   M^-1 r = D^-1 r + G (diag(G^T A G))^-1 G^T r
void ensure_gradient(Context* c, Space* s);
void ensure_nodal(Context* c, Space* s);
void ensure_entity_maps(Context* c, Space* s);
const double* gradient_diag(Context* c, System* sys, bool masked);

// batch.hip: nb driven systems (one k0 each) on one assembled system
void solve_system_batch(Context* c, System* sys, int64_t nb, const double* k0,
                        const double2* B, double2* X, const SolveOpts& o, SolveStats* st);

// sweep.hip: nb driven systems (one k0 each) solved together, `width` at a time (1, 2,
// 4 or 8 interleaved systems) by COCG with the auxiliary-space preconditioner on the
// patch operator, the iteration replayed from a hipGraph when use_graph
void solve_system_sweep(Context* c, System* sys, int64_t nb, const double* k0, const double2* B,
                        double2* X, const SolveOpts& o, int width, bool use_graph,
                        SolveStats* st);
// whether the patch operator (which the sweep needs) can be built for this system
bool sys_supports_sweep(Context* c, System* sys);

// eigen.hip
struct EigOpts {
  int ncv = 0;               // Lanczos basis size (0: see eigs_system)
  int max_restarts = 100;
  int inner_precision = 64;
  int inner_max_iter = 20000;
  double tol = 1e-10;        // relative Ritz residual of nu
  double inner_tol = 1e-11;  // inner (S - sigma T) solves
  unsigned long long seed = 12345;
  int block = 0;             // Lanczos block size 1 or 2 (0: 2 where paired inner solves run)
};
struct EigStats {
  int status = 0, nconv = 0, restarts = 0, lanczos_steps = 0, inner_failures = 0;
  int inner_stagnated = 0;
  int inner_fp64_from = 0;   // precision 32: Lanczos step from which inner solves ran in FP64
  int block = 1;             // the Lanczos block size used
  int64_t inner_iterations = 0;
  double max_ritz_residual = 0, elapsed_ms = 0;
};
// whether paired solves (SolveOpts::paired) run on this system in this precision
bool paired_solves_available(Context* c, const System* sys, int precision);
void eigs_system(Context* c, const System* sys, double sigma, int nev, const EigOpts& o,
                 double* lam_h, double* resid_h, double2* vecs_d, EigStats& st);

// assemble.hip: per boundary face, H[6 f + k] = (area / ref area) (J^-1 (I - n n^T) J^-T)_k
void face_coefs(Context* c, const Space* s, double* H_d);

// postproc.hip -- field post-processing on the boundary
// Free-space constants as the reference defines them (Consts.py:24-33):
// mu0 = pi*4*1e-7 evaluated left to right, Z0 = mu0*c0.
constexpr double kC0 = 299792458.0;
constexpr double kMu0 = M_PI * 4 * 1e-7;
constexpr double kZ0 = kMu0 * kC0;
double surface_flux(Context* c, const Space* s, const void* x, double k0, bool use_mu);
void variational_flux(Context* c, const Space* s, const void* x, double k0, double out[2]);
void ntff(Context* c, const Space* s, const void* x, double k0, bool variational, int64_t n,
          const double* theta, const double* phi, double* E_ff);

}  // namespace sucem

struct sucem_mesh { sucem::Mesh m; };
struct sucem_space { sucem::Space s; };
struct sucem_system { sucem::System s; };
struct sucem_csr { sucem::Csr a; };
