"""ctypes binding of libsucem_hip (the C-ABI declared in include/sucem_hip.h).

The library is required: importing the device API without the built library, or
running it without a HIP device, raises -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SUCEM_HIP_LIB", os.path.join(_HERE, "_native", "libsucem_hip.so"))

SUCEM_OK = 0
ERR_INVALID, ERR_HIP, ERR_OOM, ERR_OUTSIDE, ERR_MATERIAL, ERR_UNSUPPORTED, ERR_INTERNAL = (
    1, 2, 3, 4, 5, 6, 7)
MAT_S, MAT_T, MAT_S0 = 1, 2, 4
METHOD_COCG, METHOD_BICGSTAB, METHOD_GMRES = 0, 1, 2
PRECOND_NONE, PRECOND_JACOBI, PRECOND_GRADIENT, PRECOND_AUXILIARY = 0, 1, 2, 3


class SolveOpts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("precision", ctypes.c_int), ("precond", ctypes.c_int),
                ("max_iter", ctypes.c_int), ("check_every", ctypes.c_int),
                ("eliminate_pec", ctypes.c_int), ("tol", ctypes.c_double),
                ("ru_delta", ctypes.c_double), ("gmres_restart", ctypes.c_int),
                ("paired", ctypes.c_int), ("precond_shift", ctypes.c_double)]


class SolveStats(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_int),
                ("reliable_updates", ctypes.c_int), ("stagnated", ctypes.c_int),
                ("rel_residual", ctypes.c_double), ("elapsed_ms", ctypes.c_double),
                ("spmv_lo", ctypes.c_int64), ("spmv_hi", ctypes.c_int64),
                ("op_ms", ctypes.c_double), ("op_samples", ctypes.c_int64),
                ("fp64_finish", ctypes.c_int), ("batch_width", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad_"}


class EigOpts(ctypes.Structure):
    _fields_ = [("ncv", ctypes.c_int), ("max_restarts", ctypes.c_int),
                ("inner_precision", ctypes.c_int), ("inner_max_iter", ctypes.c_int),
                ("tol", ctypes.c_double), ("inner_tol", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("block", ctypes.c_int), ("pad_", ctypes.c_int)]


class EigStats(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("nconv", ctypes.c_int), ("restarts", ctypes.c_int),
                ("lanczos_steps", ctypes.c_int), ("inner_failures", ctypes.c_int),
                ("inner_stagnated", ctypes.c_int), ("inner_fp64_from", ctypes.c_int),
                ("block", ctypes.c_int), ("inner_iterations", ctypes.c_int64),
                ("max_ritz_residual", ctypes.c_double), ("elapsed_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad_"}


class SucemError(RuntimeError):
    pass


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_PP = ctypes.POINTER(ctypes.c_void_p)
_D = ctypes.c_double

# name -> (restype, argtypes); every entry point returns int status unless noted
SIGNATURES = {
    "sucem_last_error": (ctypes.c_char_p, []),
    "sucem_version": (ctypes.c_int, []),
    "sucem_ctx_create": (ctypes.c_int, [ctypes.c_int, _P, _PP]),
    "sucem_ctx_destroy": (ctypes.c_int, [_P]),
    "sucem_ctx_synchronize": (ctypes.c_int, [_P]),
    "sucem_ctx_workspace_bytes": (ctypes.c_int, [_P, _P]),
    "sucem_ctx_release_workspace": (ctypes.c_int, [_P]),
    "sucem_ctx_mem_info": (ctypes.c_int, [_P, _P, _P]),
    "sucem_mesh_create": (ctypes.c_int, [_P, _P, _I64, _P, _I64, _PP]),
    "sucem_mesh_create_box": (ctypes.c_int, [_P, _I64, _I64, _I64, _P, _P, _PP]),
    "sucem_mesh_info": (ctypes.c_int, [_P, _P, _P]),
    "sucem_mesh_export": (ctypes.c_int, [_P, _P, _P]),
    "sucem_mesh_translate": (ctypes.c_int, [_P, _P]),
    "sucem_mesh_set_materials": (ctypes.c_int, [_P, _P, _I64, _P, _P, _P]),
    "sucem_mesh_set_cell_materials": (ctypes.c_int, [_P, _P, _P]),
    "sucem_mesh_destroy": (ctypes.c_int, [_P]),
    "sucem_space_create": (ctypes.c_int, [_P, _P, ctypes.c_int, _PP]),
    "sucem_space_info": (ctypes.c_int, [_P, _P]),
    "sucem_system_patch_info": (ctypes.c_int, [_P, _P, ctypes.c_int, _P]),
    "sucem_space_export_pattern": (ctypes.c_int, [_P, _P, _P]),
    "sucem_space_export_dofmap": (ctypes.c_int, [_P, _P]),
    "sucem_space_export_entities": (ctypes.c_int, [_P, _P, _P]),
    "sucem_space_export_pec": (ctypes.c_int, [_P, _P]),
    "sucem_space_mark_pec": (ctypes.c_int, [_P, _P, _P, ctypes.c_int]),
    "sucem_system_set_spmv": (ctypes.c_int, [_P, _P, ctypes.c_int, _P]),
    "sucem_space_destroy": (ctypes.c_int, [_P]),
    "sucem_assemble": (ctypes.c_int, [_P, _P, ctypes.c_uint, _PP]),
    "sucem_system_export": (ctypes.c_int, [_P, ctypes.c_uint, _P]),
    "sucem_system_destroy": (ctypes.c_int, [_P]),
    "sucem_system_apply": (ctypes.c_int, [_P, _P, _D, ctypes.c_int, ctypes.c_int, _I64, _P, _P]),
    "sucem_system_precond_apply": (ctypes.c_int, [_P, _P, _D, _D, _D, _D, ctypes.c_int,
                                                   ctypes.c_int, _P, _P]),
    "sucem_system_apply_coef": (ctypes.c_int, [_P, _P, _D, _D, _D, ctypes.c_int, ctypes.c_int,
                                                _I64, _P, _P]),
    "sucem_system_tune": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int, _P, _P]),
    "sucem_system_export_lhs": (ctypes.c_int, [_P, _D, _P]),
    "sucem_locate_points": (ctypes.c_int, [_P, _P, _I64, _P, _P]),
    "sucem_basis_at_points": (ctypes.c_int, [_P, _P, _I64, _P, _P, _P]),
    "sucem_rhs_point_sources": (ctypes.c_int, [_P, _P, _I64, _P, _P, _P, _I64, _D, _D, _P]),
    "sucem_rhs_point_sources_noalloc": (ctypes.c_int, [_P, _P, _I64, _P, _P, _P, _I64, _D, _D,
                                                       _P]),
    "sucem_solve_opts_default": (None, [_P]),
    "sucem_solve_driven": (ctypes.c_int, [_P, _P, _D, _P, _P, _P, _P, _P, _I64]),
    "sucem_solve_driven_batch": (ctypes.c_int, [_P, _P, _I64, _P, _P, _P, _P, _P]),
    "sucem_solve_driven_sweep": (ctypes.c_int, [_P, _P, _I64, _P, _P, _P, _P, ctypes.c_int,
                                                ctypes.c_int, _P]),
    "sucem_solve_system": (ctypes.c_int, [_P, _P, _D, _D, _D, _P, _P, _P, _P, _P, _I64]),
    "sucem_eig_opts_default": (None, [_P]),
    "sucem_eigs": (ctypes.c_int, [_P, _P, _D, ctypes.c_int, _P, _P, _P, _P, _P]),
    "sucem_surface_flux": (ctypes.c_int, [_P, _P, _P, _D, ctypes.c_int, _P]),
    "sucem_variational_flux": (ctypes.c_int, [_P, _P, _P, _D, _P]),
    "sucem_ntff": (ctypes.c_int, [_P, _P, _P, _D, ctypes.c_int, _I64, _P, _P, _P]),
    "sucem_csr_create": (ctypes.c_int, [_P, _I64, _I64, _P, _P, _P, _PP]),
    "sucem_csr_apply": (ctypes.c_int, [_P, _P, _P, _P]),
    "sucem_csr_solve": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _I64]),
    "sucem_csr_destroy": (ctypes.c_int, [_P]),
}

_lib = None


def load():
    """Load the library and bind every exported signature (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libsucem_hip.so not found at %s; build it with "
                          "'python sucem-fem_amd/build_lib.py'" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


_EXC = {ERR_INVALID: ValueError, ERR_OUTSIDE: AssertionError, ERR_MATERIAL: ValueError,
        ERR_UNSUPPORTED: NotImplementedError, ERR_OOM: MemoryError}


def check(status):
    if status != SUCEM_OK:
        msg = load().sucem_last_error().decode("utf-8", "replace")
        raise _EXC.get(status, SucemError)(msg)


def call(name, *args):
    check(getattr(load(), name)(*args))


def ptr(a):
    """Address of a numpy array (C-contiguous) or a torch tensor."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return ctypes.c_void_p(a.ctypes.data)


def out_handle():
    return ctypes.c_void_p()


def f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)
