// Host-side launchers of the operator kernels (used by the Krylov layer).
#pragma once
#include "internal.hpp"
#include "kernels.hpp"

namespace sucem {

int lanes_for(int64_t n, int64_t nnz);
int op_grid(Context* c, int64_t n, int per_cu);
SpmvCfg default_cfg(Context* c, int64_t n, int64_t nnz);

// y = op(x) in precision R (float: float2 vectors / float values).  epi = 1 (or 2):
// the grid also reduces sum_r w_r y_r (conj(w_r) y_r) and sum |y_r|^2 into the block
// partials of rs; the caller finishes them with finish_reduction(c, rs, f.G).
template <typename R>
void sys_apply(Context* c, const System* sys, Coef k, const uint8_t* mask, const void* x,
               void* y, const void* w, int epi, const RedSite& rs, const SpmvCfg& f);
template <typename R>
void csr_apply(Context* c, const Csr* A, const uint8_t* mask, const void* x, void* y,
               const void* w, int epi, const RedSite& rs, const SpmvCfg& f);
template <typename R>
SpmvCfg sys_cfg(Context* c, const System* sys);
SpmvCfg csr_cfg(Context* c, const Csr* A);
template <typename R>
SpmvCfg tune_sys(Context* c, System* sys, int reps);
// matrix-free operator (matfree.hip): mf_prepare before the first mf_apply on a system
void mf_prepare(Context* c, System* sys);
template <typename R>
void mf_apply(Context* c, const System* sys, Coef k, const uint8_t* mask, const void* x, void* y,
              const void* w, int epi, const RedSite& rs, const SpmvCfg& f);

// patch operator (patch.hip) with patches of cb cells (128, 192, 256): pt_prepare before
// the first pt_apply with that cb on a system; false when the space does not suit it
bool pt_prepare(Context* c, System* sys, int cb);
template <typename R>
void pt_apply(Context* c, const System* sys, Coef k, const uint8_t* mask, const void* x, void* y,
              const void* w, int epi, const RedSite& rs, const SpmvCfg& f);
// Patch operator for COCG with the shared DOFs' partial sums left unsummed: y holds the
// rows of DOFs that lie in one patch, the partial sums stay in the system's buffer (rows
// marked by pt_qcode), and the epilogue is x.y (FP64: the sum of the element quadratic
// forms; FP32: over the finished rows and the partial sums) -- the caller's w must be x,
// zero on masked rows; rs gets pt_deferred_parts(sys, f) block partials (0: the form is
// not available for f).
int pt_deferred_parts(const System* sys, const SpmvCfg& f);
template <typename R>
void pt_apply_deferred(Context* c, const System* sys, Coef k, const uint8_t* mask, const void* x,
                       void* y, const void* w, const RedSite& rs, const SpmvCfg& f);
const int32_t* pt_qcode(const System* sys, const SpmvCfg& f);
const void* pt_partials(const System* sys);
// whether the patch kernel's LDS need in precision R fits the device
template <typename R>
bool pt_fits(const Space* s, int cb);
// the cfg of the (prepared) patch operator with a partial-sum grid of G blocks
SpmvCfg pt_cfg(const Space* s, int cb, int G);

// Second half of a grid reduction: sum the G block partials of rs and finalise its
// scalars (krylov.hip); a no-op for rs.mode == FIN_NONE.
void finish_reduction(Context* c, const RedSite& rs, int G);

// gradient-space preconditioner (hiptmair.hip)
template <typename R>
void grad_winv(Context* c, System* sys, Coef k, bool masked, bool aux, void* winv);
double gradient_weight(int order, bool aux);
// weight of the auxiliary-space preconditioner's vector nodal term for coefficients k
double nodal_weight(const Coef& k);
// vinv (nullable): the auxiliary-space form, vinv = diag(Pi^T A Pi)^-1 (3 nv) from
// nodal_winv, u a 3 nv work vector
template <typename R>
void grad_precond(Context* c, const Space* s, const void* winv, const void* dinv, const void* r,
                  void* w, void* z, const uint8_t* mask, int G, const RedSite& rs,
                  const void* vinv = nullptr, void* u = nullptr);
template <typename R>
void nodal_winv(Context* c, System* sys, Coef k, bool masked, void* vinv);
// Jacobi inverse diagonal (zero on masked rows).
template <typename R>
void sys_dinv(Context* c, const System* sys, Coef k, const uint8_t* mask, void* dinv);
template <typename R>
void csr_dinv(Context* c, const Csr* A, const uint8_t* mask, bool identity, void* dinv, int64_t n);

}  // namespace sucem
