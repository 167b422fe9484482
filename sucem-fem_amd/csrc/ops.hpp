// Host-side launchers of the operator kernels (used by the Krylov layer).
#pragma once
#include "internal.hpp"

namespace sucem {

int lanes_for(int64_t n, int64_t nnz);
int op_grid(Context* c, int64_t n);

// y = op(x) in precision R (float: float2 vectors / float values).  When `part`
// is non-null the epilogue writes part[lb] = sum op(w_r) y_r (conjugated when
// `conj`) and part[G + lb] = (sum |y_r|^2, 0) for the G blocks.
template <typename R>
void sys_apply(Context* c, const System* sys, Coef k, const uint8_t* mask, const void* x,
               void* y, const void* w, bool conj, double2* part, int G, int L);
template <typename R>
void csr_apply(Context* c, const Csr* A, const uint8_t* mask, const void* x, void* y,
               const void* w, bool conj, double2* part, int G, int L);
// Jacobi inverse diagonal (zero on masked rows).
template <typename R>
void sys_dinv(Context* c, const System* sys, Coef k, const uint8_t* mask, void* dinv);
template <typename R>
void csr_dinv(Context* c, const Csr* A, const uint8_t* mask, bool identity, void* dinv, int64_t n);

}  // namespace sucem
