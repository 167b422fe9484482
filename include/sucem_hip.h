/* C-ABI of libsucem_hip: MI355X-native Nedelec assembly, CSR operators and Krylov
 * solves behind the SUCEM:FEM problem API.
 *
 * Conventions
 *  - Every function returns an int status (SUCEM_OK = 0); on failure
 *    sucem_last_error() returns a thread-local message.
 *  - Arguments suffixed _h are host pointers; _d are device pointers on the
 *    context's device (e.g. torch.Tensor.data_ptr()).  Complex vectors are
 *    interleaved (re, im) doubles (complex128) unless stated otherwise.
 *  - Calls on one context are issued in order on the context's stream; calls
 *    returning host data synchronise that stream.
 *  - Handles own device memory; destroy them before their context.
 *
 * Reference paths below are relative to the reference repository root.
 */
#ifndef SUCEM_HIP_H
#define SUCEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SUCEM_OK = 0,
  SUCEM_ERR_INVALID = 1,       /* bad argument (reference: ValueError / AssertionError) */
  SUCEM_ERR_HIP = 2,           /* HIP runtime failure */
  SUCEM_ERR_OOM = 3,           /* device allocation failed */
  SUCEM_ERR_OUTSIDE = 4,       /* point outside the mesh (point_source.py:79 AssertionError) */
  SUCEM_ERR_MATERIAL = 5,      /* unknown material region (Materials.py:147-148 ValueError) */
  SUCEM_ERR_UNSUPPORTED = 6,   /* order / size outside what the library supports */
  SUCEM_ERR_INTERNAL = 7       /* a consistency check of the library's own data failed */
};

/* matrix selectors */
#define SUCEM_MAT_S 1u   /* curl-curl stiffness, Forms.py:79-85 */
#define SUCEM_MAT_T 2u   /* eps_r mass ("M" in the reference), Forms.py:71-77 */
#define SUCEM_MAT_S0 4u  /* first-order ABC, BoundaryConditions/ABC.py:35-58 */

/* solver methods / preconditioners */
#define SUCEM_METHOD_COCG 0
#define SUCEM_METHOD_BICGSTAB 1
/* restarted GMRES, FP64, right-preconditioned (LinalgSolvers.GMRESSolver, LinalgSolvers.py:251-260) */
#define SUCEM_METHOD_GMRES 2
#define SUCEM_PRECOND_NONE 0
#define SUCEM_PRECOND_JACOBI 1
/* Jacobi plus the gradient-space correction G diag(G^T A G)^-1 G^T (driven / shifted
 * Nedelec systems, COCG only) */
#define SUCEM_PRECOND_GRADIENT 2
/* Auxiliary-space (Hiptmair-Xu) form: the gradient term plus the vector nodal term
 * Pi diag(Pi^T A Pi)^-1 Pi^T, Pi the nodal interpolation of piecewise-linear vector fields
 * (COCG only) */
#define SUCEM_PRECOND_AUXILIARY 3

typedef struct sucem_ctx sucem_ctx;
typedef struct sucem_mesh sucem_mesh;
typedef struct sucem_space sucem_space;
typedef struct sucem_system sucem_system;
typedef struct sucem_csr sucem_csr;

typedef struct sucem_solve_opts {
  int method;         /* SUCEM_METHOD_* */
  int precision;      /* 64: FP64 throughout; 32: FP32 matrix+vectors, FP64 reliable updates */
  int precond;        /* SUCEM_PRECOND_* */
  int max_iter;
  int check_every;    /* host polls the residual every this many iterations */
  int eliminate_pec;  /* 1: restrict to non-PEC DOFs (rows/cols of PEC DOFs removed) */
  double tol;         /* relative residual ||b - A x|| / ||b|| (FP64) */
  double ru_delta;    /* FP32: reliable update when the residual fell by this factor */
  int gmres_restart;  /* GMRES: Arnoldi steps per cycle (SciPy's default 20) */
  int paired;         /* 1: b and x hold two real systems, one in the real and one in the
                         imaginary parts, solved as two independent real COCG recursions that
                         share every operator and preconditioner pass (sucem_solve_system
                         with c0 = 0 and precond_shift 0, gradient or auxiliary-space
                         preconditioner, patch operator; the residual reported is the larger
                         of the two relative residuals).  No counterpart in the reference:
                         its real eigen pencils are solved one right-hand side at a time. */
  double precond_shift; /* complex shift beta of the operator the Jacobi / gradient /
                           auxiliary diagonals are built from: cs S + ct (1 - i beta) T +
                           i c0 S0 (the solved operator is unchanged).  < 0 (the default):
                           0.25 for the driven entry points (fewer COCG iterations on driven
                           Helmholtz systems), 0 for sucem_solve_system and generic matrices
                           (the shifted cavity operator S - sigma T converges slower with it) */
} sucem_solve_opts;

typedef struct sucem_eig_opts {
  int ncv;              /* Lanczos basis size (0: max(2 nev + 1, 20) for block size 1,
                           max(2 nev + 2, nev + 20) for block size 2) */
  int max_restarts;
  int inner_precision;  /* 64: FP64 inner solves; 32: FP32 refinement cycles (FP32 COCG
                           recursions restarted from the FP64 residual, x in FP64) */
  int inner_max_iter;
  double tol;           /* relative Ritz residual of nu = 1/(k^2 - sigma) */
  double inner_tol;     /* relative residual of the inner (S - sigma T) solves */
  uint64_t seed;        /* start vector */
  int block;            /* Lanczos block size: 1, 2 (each pair of inner solves runs as one
                           paired solve, see sucem_solve_opts.paired), 0 (default): 2 where the
                           paired solve is available (patch operator), else 1 */
  int pad_;
} sucem_eig_opts;

typedef struct sucem_eig_stats {
  int status;           /* 0 converged, 1 restart limit reached */
  int nconv;
  int restarts;
  int lanczos_steps;    /* Lanczos basis vectors generated (inner solves x block size) */
  int inner_failures;   /* inner solves that hit inner_max_iter or broke down */
  int inner_stagnated;  /* inner solves stopped at the floor of their true residual */
  int inner_fp64_from;  /* inner_precision 32: the Lanczos vector (counted as in
                           lanczos_steps) from which the inner solves ran in FP64 because two
                           solves in a row had their FP32 cycles fail and were finished in FP64
                           (0: all inner solves ran FP32 cycles) */
  int block;            /* the Lanczos block size used (lanczos_steps counts basis vectors:
                           lanczos_steps / block inner solves) */
  int64_t inner_iterations;  /* COCG iterations over all inner solves (a paired solve's
                                once) */
  double max_ritz_residual;
  double elapsed_ms;
} sucem_eig_stats;

typedef struct sucem_solve_stats {
  int status;           /* 0 converged, >0 iteration limit or stagnation, <0 breakdown
                           (scipy 'info') */
  int iterations;
  int reliable_updates;
  int stagnated;        /* 1: stopped because three residual replacements in a row did not
                           halve the true residual (its FP64 floor lies above tol) */
  double rel_residual;  /* final FP64 ||b - A x|| / ||b|| */
  double elapsed_ms;
  int64_t spmv_lo;      /* SpMVs in the working precision */
  int64_t spmv_hi;      /* FP64 residual SpMVs */
  double op_ms;         /* mean device time (ms) of one working-precision operator application
                           (its kernels, with the fused dot-product epilogue), measured with
                           HIP events on the solve's stream at one iteration per check interval */
  int64_t op_samples;   /* number of such samples (0: none taken) */
  int fp64_finish;      /* precision 32: 1 when the FP32 cycles stopped lowering the true
                           residual (or diverged) and the solve was finished in FP64 from
                           their accepted solution */
  int batch_width;      /* systems solved together in one group (batched sweep; 1 otherwise) */
} sucem_solve_stats;

const char* sucem_last_error(void);
int sucem_version(void);

/* Context: device and the stream all work is issued on (NULL = the null stream).  A
 * context keeps the solvers' work vectors resident (a workspace that grows to the largest
 * solve run on it), so repeated solves allocate no device memory; calls on one context
 * must come from one host thread at a time. */
int sucem_ctx_create(int device, void* hip_stream, sucem_ctx** out);
int sucem_ctx_destroy(sucem_ctx* ctx);
int sucem_ctx_synchronize(sucem_ctx* ctx);
/* Device bytes held by the context's workspace; release frees it (the next solve
 * allocates again).  (No reference counterpart: memory accounting for serving.) */
int sucem_ctx_workspace_bytes(sucem_ctx* ctx, int64_t* bytes_h);
int sucem_ctx_release_workspace(sucem_ctx* ctx);
/* Free and total device memory of the context's device (hipMemGetInfo).  (No reference
 * counterpart: admission of requests by device memory.) */
int sucem_ctx_mem_info(sucem_ctx* ctx, int64_t* free_h, int64_t* total_h);

/* ---- mesh (replaces the DOLFIN Mesh handed to EMProblem.set_mesh, EMProblem.py:49-54) ---- */
/* Cells are sorted per row on the device (DOLFIN mesh ordering). */
int sucem_mesh_create(sucem_ctx* ctx, const double* xyz_h, int64_t nv, const int64_t* cells_h,
                      int64_t nc, sucem_mesh** out);
/* DOLFIN UnitCube(nx,ny,nz) Kuhn mesh, x -> x*scale + shift per axis, generated on the
 * device (Utilities/MeshGenerators.py:26-53 without the centring step). */
int sucem_mesh_create_box(sucem_ctx* ctx, int64_t nx, int64_t ny, int64_t nz,
                          const double* scale3_h, const double* shift3_h, sucem_mesh** out);
int sucem_mesh_info(const sucem_mesh* m, int64_t* nv, int64_t* nc);
int sucem_mesh_export(sucem_mesh* m, double* xyz_h, int64_t* cells_h);
/* Translate every vertex by -shift3 (the centring step of get_centred_cube). */
int sucem_mesh_translate(sucem_mesh* m, const double* shift3_h);
/* Per-cell eps_r and 1/mu_r from a region id per cell (NULL: all region 0) and a
 * region table (MaterialFunctionFactory.get_material_functions, Materials.py:118-150).
 * Unknown region -> SUCEM_ERR_MATERIAL. */
int sucem_mesh_set_materials(sucem_mesh* m, const int64_t* region_h, int64_t nregions,
                             const int64_t* region_ids_h, const double* eps_r_h,
                             const double* mu_r_h);
/* Per-cell eps_r and 1/mu_r given directly (nc doubles each, finite; zero weights allowed):
 * the cellwise-constant material functions of CalcEMFunctional.set_epsr_function /
 * set_mur_function (PostProcessing/PostProcessing.py:103-110), with 1/mu_r = eps_r = 0
 * outside a cell-domain mark (set_cell_domains, PostProcessing.py:93-101). */
int sucem_mesh_set_cell_materials(sucem_mesh* m, const double* eps_r_h, const double* inv_mu_r_h);
int sucem_mesh_destroy(sucem_mesh* m);

/* ---- function space (replaces dolfin.FunctionSpace(mesh, "Nedelec 1st kind H(curl)", order),
 *      EMProblem.py:100-105): first-encounter numbering, DOF map, CSR pattern, PEC DOFs ---- */
int sucem_space_create(sucem_ctx* ctx, sucem_mesh* m, int order, sucem_space** out);
/* info[0..7] = ndofs, nnz, nedges, nfaces, nboundary_faces, dofs_per_cell, nnz_abc, npec */
int sucem_space_info(const sucem_space* s, int64_t* info8_h);
/* CSR pattern as the uBLAS export gives it (Utilities/Converters.py:27-44): indptr[N+1],
 * indices[nnz] (sorted columns). */
int sucem_space_export_pattern(sucem_space* s, int64_t* indptr_h, int32_t* indices_h);
int sucem_space_export_dofmap(sucem_space* s, int64_t* cell_dofs_h);
int sucem_space_export_entities(sucem_space* s, int64_t* edges_h, int64_t* faces_h);
/* PEC (essential) DOFs: by default the closure of every boundary facet, the PEC walls of
 * essential.py:86-99. */
int sucem_space_export_pec(sucem_space* s, uint8_t* mask_h);
/* Set the PEC DOFs to the closure of the facets flagged in facet_flag_h (nfaces bytes, in
 * the space's facet numbering; null = every boundary facet), replacing the current set or,
 * with union_mode != 0, adding to it: the DirichletBC(V, 0, mesh_function, region) of
 * essential.py:65-81 with facet_flag_h[f] = (mesh_function[f] == region). */
int sucem_space_mark_pec(sucem_ctx* ctx, sucem_space* s, const uint8_t* facet_flag_h,
                         int union_mode);
int sucem_space_destroy(sucem_space* s);

/* ---- assembly (replaces SystemMatrices.calc_system_matrices, SystemMatrices.py:46-58) ---- */
int sucem_assemble(sucem_ctx* ctx, sucem_space* s, unsigned which, sucem_system** out);
/* Values of one matrix (SUCEM_MAT_S, _T or _S0) on the full pattern, explicit zeros kept. */
int sucem_system_export(sucem_system* sys, unsigned which_one, double* values_h);
int sucem_system_destroy(sucem_system* sys);

/* ---- driven operator (replaces DrivenProblemABC.get_LHS_matrix, EMDrivenProblem.py:67-72):
 *      y = (S - k0^2 T + i k0 S0) x for nvec complex vectors, precision 64 (complex128 x/y)
 *      or 32 (complex64 x/y and FP32 matrix values). ---- */
int sucem_system_apply(sucem_ctx* ctx, sucem_system* sys, double k0, int precision,
                       int eliminate_pec, int64_t nvec, const void* x_d, void* y_d);
/* General combination y = (cs S + ct T + i c0 S0) x (e.g. T x for M*x, S - sigma T for the
 * eigen shift of EMVectorWaveEigenproblem.py:74-76). */
int sucem_system_apply_coef(sucem_ctx* ctx, sucem_system* sys, double cs, double ct, double c0,
                            int precision, int eliminate_pec, int64_t nvec, const void* x_d,
                            void* y_d);
/* Autotune the fused operator for precision 32 or 64 by timing every candidate with HIP
 * events: the assembled CSR operator over (lanes per row, loads per lane, blocks), the
 * matrix-free operator (element products + row gather) over its gather grid and the patch
 * operator (element products summed per patch of cells in LDS + shared-DOF sums) over the
 * grid of its shared-DOF kernel.  The fastest is kept on the system and used by apply and
 * solve.  cfg4_h (nullable) receives {matrix_free, lanes, loads, blocks}, ms_h the time
 * per application.  (No reference counterpart: launch tuning of the device operator.) */
int sucem_system_tune(sucem_ctx* ctx, sucem_system* sys, int precision, int reps, int* cfg4_h,
                      double* ms_h);
/* Set the operator explicitly: cfg4_h = {matrix_free (0 assembled, 1 matrix-free per cell,
 * 2 matrix-free per patch), lanes (4, 8, 16, 32; matrix-free: lanes per row of its gather,
 * 1 or 4, or 0 for one row per thread staged through LDS; patch: ignored), loads (1, 2, 4;
 * matrix-free: 1 = element rows in one pass, 2 = half the rows at a time; patch: cells
 * per patch, 128, 192 or 256 (order 1: 256)),
 * blocks (multiple of 8; patch: grid of the shared-DOF kernel)}.
 * Calls on one system must not overlap (the matrix-free element buffer is per system). */
int sucem_system_set_spmv(sucem_ctx* ctx, sucem_system* sys, int precision, const int* cfg4_h);
/* Sizes of the patch operator's maps for patches of cb cells (built if needed):
 * info6_h = {patches, patch DOFs, DOFs shared by patches, partial sums, largest LDS
 * entry count of a patch, largest (boundary face, row) count of a patch}; for the
 * operator's byte count.  (No reference counterpart.) */
int sucem_system_patch_info(sucem_ctx* ctx, sucem_system* sys, int cb, int64_t* info6_h);
/* Materialised A values (complex128 interleaved) on the full pattern. */
int sucem_system_export_lhs(sucem_system* sys, double k0, double* values_h);

/* ---- sources (replaces calc_pointsource_contrib, Sources/point_source.py:49-94) ---- */
/* Lowest-index cell containing each point; SUCEM_ERR_OUTSIDE if a point is in none. */
int sucem_locate_points(sucem_ctx* ctx, sucem_space* s, int64_t n, const double* pts_h,
                        int64_t* cells_h);
/* Physical basis values N_i(x) (n x dofs_per_cell x 3) and the cell DOFs (n x dofs_per_cell). */
int sucem_basis_at_points(sucem_ctx* ctx, sucem_space* s, int64_t n, const double* pts_h,
                          double* vals_h, int64_t* dofs_h);
/* b[r] += scale * sum_{s: rhs(s)=r} N(x_s) . p_s  for nrhs complex vectors of length N,
 * sources added in input order (DrivenProblemABC.get_RHS, EMDrivenProblem.py:74-80, with
 * scale = -i k0 Z0).  pos_h: nsrc x 3, val_h: nsrc x 3 complex, rhs_h: nsrc indices. */
int sucem_rhs_point_sources(sucem_ctx* ctx, sucem_space* s, int64_t nsrc, const double* pos_h,
                            const double* val_h, const int64_t* rhs_h, int64_t nrhs,
                            double scale_re, double scale_im, double* b_d);

/* Allocation-free variant of sucem_rhs_point_sources (same arguments, same result).  No
 * counterpart in the reference: it is for callers that cannot allocate device memory
 * (serving loops, code that forbids hipMalloc).  At most 4096 sources per call
 * (SUCEM_ERR_INVALID beyond); their positions, moments, indices and cells are staged in a
 * fixed array of device memory that the library holds on each GPU until the process ends
 * (~352 KB), so the call makes no hipMalloc, hipFree or pinned-host allocation.  Calls
 * that overlap on one GPU (from several contexts or threads) race on that array. */
int sucem_rhs_point_sources_noalloc(sucem_ctx* ctx, sucem_space* s, int64_t nsrc,
                                    const double* pos_h, const double* val_h,
                                    const int64_t* rhs_h, int64_t nrhs, double scale_re,
                                    double scale_im, double* b_d);

/* ---- solves (replaces LinalgSolvers.BiCGStabSolver/solve_sparse_system,
 *      Utilities/LinalgSolvers.py:239-248,308-322, on the driven system) ---- */
void sucem_solve_opts_default(sucem_solve_opts* o);
int sucem_solve_driven(sucem_ctx* ctx, sucem_system* sys, double k0, const double* b_d,
                       double* x_d, const sucem_solve_opts* opts, sucem_solve_stats* stats,
                       double* res_hist_h, int64_t hist_cap);

/* nb driven systems on one assembled system, system j at wavenumber k0_h[j]:
 * (S - k0_j^2 T + i k0_j S0) x_j = b_j.  B_d, X_d: nb x N complex128 (system-major).  COCG
 * with SUCEM_PRECOND_AUXILIARY (the default, opts = NULL) runs the batched sweep below (8
 * systems at a time, hipGraph); FP64 COCG with SUCEM_PRECOND_GRADIENT reads the assembled
 * matrix once per iteration for up to 16 interleaved systems; other options solve the
 * systems one after another.  stats_h (nullable): nb entries.
 * (Frequency sweeps of the reference re-form and re-solve the LHS per frequency,
 * sandbox/patch/patch_driver.py:64-129.) */
int sucem_solve_driven_batch(sucem_ctx* ctx, sucem_system* sys, int64_t nb, const double* k0_h,
                             const double* B_d, double* X_d, const sucem_solve_opts* opts,
                             sucem_solve_stats* stats_h);
/* The batched sweep: the nb systems are solved `width` (1, 2, 4 or 8; 0: autotuned -- the
 * first systems are solved in timed groups of 2, 4 and 8 and the width with the least time
 * per system-iteration takes the rest; stats_h[j].batch_width reports it) at a time, each
 * group as that many COCG recursions (auxiliary-space preconditioner, precision 64 or 32
 * with FP64 residual replacement per system) over vectors interleaved per DOF, whose
 * operator reads each patch's maps and cell coefficients once for the whole group; with
 * use_graph the check interval's iterations are captured once in a hipGraph and replayed.
 * Results do not depend on width or use_graph beyond rounding of the FP32 recursions.
 * (Replaces the per-frequency / per-source loops of sandbox/patch/patch_driver.py:106-129
 * and Sources/point_source.py:49-94.) */
int sucem_solve_driven_sweep(sucem_ctx* ctx, sucem_system* sys, int64_t nb, const double* k0_h,
                             const double* B_d, double* X_d, const sucem_solve_opts* opts,
                             int width, int use_graph, sucem_solve_stats* stats_h);
/* z = M^-1 r, the auxiliary-space preconditioner of (cs S + ct T + i c0 S0) that COCG applies
 * (Jacobi + gradient + vector nodal terms, diagonals built with the complex shift `shift`),
 * once, in precision 32 (r, z complex64) or 64 (complex128); PEC rows zero with
 * eliminate_pec.  No counterpart in the reference (its preconditioners are SciPy's): for
 * testing the preconditioner kernels against a restatement of M^-1. */
int sucem_system_precond_apply(sucem_ctx* ctx, sucem_system* sys, double cs, double ct,
                               double c0, double shift, int precision, int eliminate_pec,
                               const void* r_d, void* z_d);
/* Solve (cs S + ct T + i c0 S0) x = b (e.g. the shifted S - sigma T of the eigen solver). */
int sucem_solve_system(sucem_ctx* ctx, sucem_system* sys, double cs, double ct, double c0,
                       const double* b_d, double* x_d, const sucem_solve_opts* opts,
                       sucem_solve_stats* stats, double* res_hist_h, int64_t hist_cap);

/* ---- cavity eigenproblem (replaces DefaultEigenSolver.solve_problem,
 *      EMVectorWaveEigenproblem.py:61-105): the nev lowest k^2 > sigma of the PEC-eliminated
 *      pencil (S, T), ascending, returned as k^2 directly (the reference returns
 *      nu = 1/(k^2 - sigma)).  Only modes above sigma: the reference's ARPACK shift-invert
 *      (which='LM') returns the nev modes nearest sigma on either side, which differs when
 *      sigma lies between modes; place sigma below the wanted band (the reference drivers
 *      use 1.1, below the lowest cavity mode).  resid_h (nullable): ||S x - k^2 T x|| /
 *      (k^2 ||T x||);
 *      vecs_d (nullable): nev x N complex128 T-normalised modes (zero on PEC DOFs). ---- */
void sucem_eig_opts_default(sucem_eig_opts* o);
int sucem_eigs(sucem_ctx* ctx, sucem_system* sys, double sigma, int nev,
               const sucem_eig_opts* opts, double* k2_h, double* resid_h, double* vecs_d,
               sucem_eig_stats* stats);

/* ---- post-processing of a solution x_d (complex128[N] on the device) ----
 * Real power flux through the mesh boundary,
 *   (1/(k0 Z0)) int n . (E_r x (-curl E_i / mu_r) + E_i x (curl E_r / mu_r)) dS,
 * replacing SurfaceFlux.calc_flux (PostProcessing/power_flux.py:52-63).  use_mu = 0 takes
 * mu_r = 1 (the reference's default), 1 the mesh's per-cell mu_r. */
int sucem_surface_flux(sucem_ctx* ctx, sucem_space* V, const double* x_d, double k0, int use_mu,
                       double* flux);
/* Variational power flux, replacing VariationalSurfaceFlux.calc_flux
 * (power_flux.py:65-133): conj of the free-space functional <curl E, curl g> - k0^2 <E, g>
 * over the cells touching the boundary through an edge, g = j conj(E)/(k0 Z0) on the boundary
 * DOFs.  flux: complex (2 doubles). */
int sucem_variational_flux(sucem_ctx* ctx, sucem_space* V, const double* x_d, double k0,
                           double* flux);
/* Far field at unit radius, (E_theta, E_phi) per direction (theta_h, phi_h in radians):
 * NTFF.calc_pt of PostProcessing/surface_ntff.py:125-155 (variational = 0) or
 * PostProcessing/variational_ntff.py:91-131 (variational = 1).  E_ff_h: complex128[n][2]. */
int sucem_ntff(sucem_ctx* ctx, sucem_space* V, const double* x_d, double k0, int variational,
               int64_t n, const double* theta_h, const double* phi_h, double* E_ff_h);

/* ---- generic complex CSR (LinalgSolvers on a user scipy matrix, LinalgSolvers.py:30-322) ---- */
int sucem_csr_create(sucem_ctx* ctx, int64_t n, int64_t nnz, const int64_t* indptr_h,
                     const int32_t* indices_h, const double* values_h, sucem_csr** out);
int sucem_csr_apply(sucem_ctx* ctx, sucem_csr* A, const double* x_d, double* y_d);
int sucem_csr_solve(sucem_ctx* ctx, sucem_csr* A, const double* b_d, double* x_d,
                    const sucem_solve_opts* opts, sucem_solve_stats* stats, double* res_hist_h,
                    int64_t hist_cap);
int sucem_csr_destroy(sucem_csr* A);

#ifdef __cplusplus
}
#endif

#endif /* SUCEM_HIP_H */
