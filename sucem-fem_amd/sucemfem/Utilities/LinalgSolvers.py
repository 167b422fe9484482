"""Sparse linear solvers on the device (sucemfem/Utilities/LinalgSolvers.py:30-322).

Same classes, logging data and call pattern as the reference.  The Krylov
iterations run in libsucem_hip: a matrix produced by
``DrivenProblemABC.get_LHS_matrix`` is solved with the fused S/T/S0 operator,
any other matrix is uploaded as a complex CSR.  Preconditioners: ``None``,
``'diagonal'`` (Jacobi), ``'gradient'`` (Jacobi plus the discrete-gradient
correction, COCG on a driven system only), ``'auxiliary'`` (the gradient form plus the
vector nodal correction, COCG on a driven system only) and ``'ilu'``.  The device has no
incomplete-LU factorisation: ``'ilu'`` selects the auxiliary-space preconditioner for a
driven system solved by COCG and Jacobi otherwise, with a warning.  Tolerances default to
SciPy's (relative residual 1e-5, at most 10*N iterations) as in the reference,
which passes none.
"""
import ctypes
import time
import warnings

import numpy as np
import scipy.sparse

from .. import _hip
from ..runtime import Handle, get_context


def _device_opts(method, precond, tol, maxiter, precision, check_every):
    o = _hip.SolveOpts()
    _hip.load().sucem_solve_opts_default(ctypes.byref(o))
    o.method = method
    o.precond = precond
    o.tol = float(tol)
    o.max_iter = int(maxiter)
    o.precision = int(precision)
    o.check_every = int(check_every)
    return o


def _with_restart(o, restart):
    o.gmres_restart = int(restart)
    return o


class DeviceCSR(object):
    """A complex CSR matrix uploaded to the device."""

    def __init__(self, A, ctx=None):
        self.ctx = ctx or get_context()
        A = scipy.sparse.csr_matrix(A, dtype=np.complex128)
        A.sort_indices()
        self.shape = A.shape
        indptr = np.ascontiguousarray(A.indptr, np.int64)
        indices = np.ascontiguousarray(A.indices, np.int32)
        data = np.ascontiguousarray(A.data, np.complex128)
        self.handle = Handle(self.ctx, "sucem_csr_destroy")
        _hip.call("sucem_csr_create", self.ctx.handle, A.shape[0], A.nnz, _hip.ptr(indptr),
                  _hip.ptr(indices), _hip.ptr(data), ctypes.byref(self.handle.h))

    def apply(self, x_dev, y_dev):
        _hip.call("sucem_csr_apply", self.ctx.handle, self.handle.h, _hip.ptr(x_dev), _hip.ptr(y_dev))


class SystemSolverBase(object):
    """Logging base class; subclasses choose the Krylov method."""
    method = _hip.METHOD_BICGSTAB
    default_tol = 1e-5

    def __init__(self, A, preconditioner_type=None, tol=None, maxiter=None, precision=64,
                 check_every=8):
        self._logging_data = {'time': [], 'id': [], 'res': []}
        self._user_callbacks = []
        self._timestamp('init')
        self._A = A
        self.tol = self.default_tol if tol is None else tol
        self.maxiter = maxiter
        self.precision = precision
        self.check_every = check_every
        self.stats = None
        self.set_preconditioner(preconditioner_type)
        self._callback_count = 0

    def add_user_callback(self, cb):
        self._user_callbacks.append(cb)

    def _timestamp(self, id, res=np.nan):
        self._logging_data['time'].append(time.time())
        self._logging_data['id'].append(id)
        self._logging_data['res'].append(res)
        for cb in self._user_callbacks:
            cb(self, res)

    def set_b(self, b):
        self._b = b

    def set_preconditioner(self, M_type):
        self._timestamp('preconditioner::start')
        if M_type is None:
            self._precond = _hip.PRECOND_NONE
        elif M_type.lower() == 'diagonal':
            self._precond = _hip.PRECOND_JACOBI
        elif M_type.lower() == 'gradient':
            self._precond = _hip.PRECOND_GRADIENT
        elif M_type.lower() == 'auxiliary':
            self._precond = _hip.PRECOND_AUXILIARY
        elif M_type.lower() == 'ilu':
            driven = getattr(self._A, "device_operator", None) is not None
            if driven and self.method in (_hip.METHOD_COCG, _hip.METHOD_GMRES):
                warnings.warn("incomplete LU is not available on the device; using the "
                              "auxiliary-space preconditioner")
                self._precond = _hip.PRECOND_AUXILIARY
            else:
                warnings.warn("incomplete LU is not available on the device; using Jacobi "
                              "('diagonal')")
                self._precond = _hip.PRECOND_JACOBI
        else:
            warnings.warn("Preconditioner type '%s' not recognised. Not using preconditioner." % M_type)
            self._precond = _hip.PRECOND_NONE
        self._M_type = M_type
        self._timestamp('preconditioner::end')

    def _call_solver(self):
        A, b = self._A, np.asarray(self._b)
        shape = b.shape
        b = np.ascontiguousarray(b.reshape(-1), np.complex128)
        n = A.shape[0]
        maxiter = self.maxiter or 10 * n
        opts = _device_opts(self.method, self._precond, self.tol, maxiter, self.precision,
                            self.check_every)
        if self.method == _hip.METHOD_GMRES:
            opts = _with_restart(opts, self.restart)
        stats = _hip.SolveStats()
        hist = np.zeros(maxiter, np.float64)
        op = getattr(A, "device_operator", None)
        if op is not None:
            system, k0 = op[0], op[1]
            pec_d = op[3] if len(op) > 3 else None
            ctx = system.ctx
            b_d = ctx.to_device(b)
            x_d = ctx.tensor(n)
            xp_d = None
            if pec_d is not None:
                # PEC rows are pec_d * x_i = b_i (DirichletBC rows); the other rows keep
                # their couplings to the PEC columns, so b_i / pec_d on the PEC DOFs is moved
                # to the right-hand side and the solve runs on the free DOFs only
                mask = system.function_space.pec_mask()
                opts.eliminate_pec = 1
                if np.any(b[mask] != 0):
                    x_pec = np.zeros(n, np.complex128)
                    x_pec[mask] = b[mask] / pec_d
                    xp_d = ctx.to_device(x_pec)
                    y_d = ctx.tensor(n)
                    _hip.call("sucem_system_apply", ctx.handle, system.handle.h, float(k0), 64,
                              0, 1, _hip.ptr(xp_d), _hip.ptr(y_d))
                    b_d = b_d - y_d
            _hip.call("sucem_solve_driven", ctx.handle, system.handle.h, float(k0), _hip.ptr(b_d),
                      _hip.ptr(x_d), ctypes.byref(opts), ctypes.byref(stats), _hip.ptr(hist),
                      maxiter)
            if xp_d is not None:
                x_d = x_d + xp_d
        else:
            dev = DeviceCSR(A)
            ctx = dev.ctx
            b_d = ctx.to_device(b)
            x_d = ctx.tensor(n)
            _hip.call("sucem_csr_solve", ctx.handle, dev.handle.h, _hip.ptr(b_d), _hip.ptr(x_d),
                      ctypes.byref(opts), ctypes.byref(stats), _hip.ptr(hist), maxiter)
        self.stats = stats.as_dict()
        nb = np.linalg.norm(b)
        for k in range(stats.iterations):
            self._callback_count += 1
            self._timestamp(self._callback_count, res=hist[k] * nb)
        x = x_d.cpu().numpy()
        if np.isrealobj(self._b) and np.isrealobj(A.data if hasattr(A, "data") else 0):
            x = x.real.copy()
        info = 0 if stats.status == 0 else (stats.iterations if stats.status > 0 else -1)
        return x.reshape(shape) if len(shape) == 1 else x, info

    def solve(self, b):
        """Solve A x = b; non-convergence is reported through ``self.stats``, as the
        reference ignores SciPy's info flag (LinalgSolvers.py:142-145)."""
        self._timestamp('solve::begin')
        self.set_b(b)
        x, info = self._call_solver()
        self.info = info
        self._timestamp('solve::end')
        return x

    def get_logging_data(self):
        return self._logging_data

    def _get_time_for_id(self, id):
        return self._logging_data['time'][self._logging_data['id'].index(id)]

    def _get_elapsed_time(self, id0, id1, force_total=False):
        if force_total:
            return self._logging_data['time'][-1] - self._logging_data['time'][0]
        return self._get_time_for_id(id1) - self._get_time_for_id(id0)

    def print_logging_data(self):
        for k in self._logging_data:
            print(k)
            print(self._logging_data[k])

    def print_timing_info(self):
        print('solve time:', self._get_elapsed_time('solve::begin', 'solve::end'))
        print('preconditioner time:',
              self._get_elapsed_time('preconditioner::start', 'preconditioner::end'))
        print('total time:', self._get_elapsed_time(None, None, True))


class BiCGStabSolver(SystemSolverBase):
    """Preconditioned BiCGStab (LinalgSolvers.py:239-248)."""
    method = _hip.METHOD_BICGSTAB


class COCGSolver(SystemSolverBase):
    """Conjugate-orthogonal CG for complex-symmetric A (one SpMV per iteration)."""
    method = _hip.METHOD_COCG


class GMRESSolver(SystemSolverBase):
    """Restarted GMRES (LinalgSolvers.py:251-260; SciPy's restart length 20), FP64 with
    right preconditioning, so the residual it monitors is the true ||b - A x||."""
    method = _hip.METHOD_GMRES
    restart = 20

    def __init__(self, A, preconditioner_type=None, restart=20, **kw):
        self.restart = int(restart)
        SystemSolverBase.__init__(self, A, preconditioner_type, **kw)


class UMFPACKSolver(SystemSolverBase):
    """Stand-in for the direct UMFPACK solve (LinalgSolvers.py:262-277).

    There is no sparse direct factorisation on the device; the system is solved
    iteratively to a relative residual of 1e-12 instead: COCG with the auxiliary-space
    preconditioner for a driven operator (complex symmetric), Jacobi BiCGStab for any
    other matrix.  A direct solve always returns the solution, so a solve that misses
    the tolerance is reported with a ``RuntimeWarning`` (the iterate is still returned,
    and ``self.stats`` holds the final relative residual).
    """
    method = _hip.METHOD_BICGSTAB
    default_tol = 1e-12

    def __init__(self, A, preconditioner_type=None, **kw):
        if getattr(A, "device_operator", None) is not None:
            self.method = _hip.METHOD_COCG
            preconditioner_type = preconditioner_type or 'auxiliary'
        SystemSolverBase.__init__(self, A, preconditioner_type or 'diagonal', **kw)

    def solve(self, b):
        x = SystemSolverBase.solve(self, b)
        if self.stats is not None and self.stats['status'] != 0:
            warnings.warn("UMFPACKSolver stand-in: the iterative solve stopped at relative "
                          "residual %.3e (status %d) above the tolerance %.1e; the iterate is "
                          "not the direct solution"
                          % (self.stats['rel_residual'], self.stats['status'], self.tol),
                          RuntimeWarning)
        return x


def calculate_residual(A, x, b):
    """||A x - b||_2 with b reshaped to x (LinalgSolvers.py:297-305)."""
    return np.linalg.norm(A * x - np.asarray(b).reshape(np.shape(x)))


def solve_sparse_system(A, b, preconditioner_type='ilu', **kw):
    """BiCGStab solve of A x = b (LinalgSolvers.py:308-322)."""
    solver = BiCGStabSolver(A, preconditioner_type, **kw)
    return solver.solve(b)
