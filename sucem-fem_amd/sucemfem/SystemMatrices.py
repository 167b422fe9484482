"""Assembly of the named system matrices (sucemfem/SystemMatrices.py:26-58).

``calc_system_matrices`` assembles every requested form in one device pass
(``sucem_assemble``: S, T and S0 share one CSR pattern and one gather map) and
applies the essential boundary conditions.  Each entry of the returned dict is a
:class:`DeviceMatrix` (or ``None`` for a ``NullForm``); ``to_scipy()`` exports it
as the CSR matrix the reference obtains from DOLFIN (Utilities/Converters.py:27-44).
"""
from __future__ import division

import ctypes

import numpy as np
import scipy.sparse

from . import Forms
from . import _hip
from .runtime import Handle

_BITS = {'S': _hip.MAT_S, 'T': _hip.MAT_T, 'S0': _hip.MAT_S0}


class AssembledSystem(object):
    """Device-resident S / T / S0 values of one function space."""

    def __init__(self, function_space, which):
        self.function_space = function_space
        self.ctx = function_space.ctx
        self.which = which
        self.handle = Handle(self.ctx, "sucem_system_destroy")
        _hip.call("sucem_assemble", self.ctx.handle, function_space.handle.h, which,
                  ctypes.byref(self.handle.h))

    def values(self, name):
        nnz = int(self.function_space.info()[1])
        out = np.empty(nnz, np.float64)
        _hip.call("sucem_system_export", self.handle.h, _BITS[name], _hip.ptr(out))
        return out

    def lhs_values(self, k0):
        nnz = int(self.function_space.info()[1])
        out = np.empty(nnz, np.complex128)
        _hip.call("sucem_system_export_lhs", self.handle.h, float(k0), _hip.ptr(out))
        return out


class DeviceMatrix(object):
    """One assembled matrix ('S', 'T' or 'S0') of an :class:`AssembledSystem`."""

    def __init__(self, system, name):
        self.system = system
        self.name = name
        self.pec_constrained = False

    def size(self, dim):
        return self.system.function_space.dim()

    @property
    def shape(self):
        n = self.system.function_space.dim()
        return (n, n)

    def constrain_pec(self):
        """Record DOLFIN ``DirichletBC.apply`` on this matrix (zero PEC rows, unit diagonal)."""
        self.pec_constrained = True

    def to_scipy(self, dtype=None):
        indptr, indices = self.system.function_space.pattern()
        vals = self.system.values(self.name)
        A = scipy.sparse.csr_matrix((vals, indices.astype(np.int32), indptr.astype(np.int32)),
                                    shape=self.shape, dtype=dtype)
        if self.pec_constrained:
            mask = self.system.function_space.pec_mask()
            rows = np.repeat(np.arange(self.shape[0]), np.diff(indptr))
            kill = mask[rows]
            A.data[kill] = 0
            diag = kill & (indices == rows)
            A.data[diag] = 1
        return A

    def array(self):
        return self.to_scipy().toarray()


class SystemMatrices(object):
    MatrixClass = DeviceMatrix

    def set_matrix_class(self, matrix_class):
        self.MatrixClass = matrix_class

    def set_matrix_forms(self, matrix_forms):
        self.matrix_forms = matrix_forms

    def set_boundary_conditions(self, boundary_conditions):
        self.boundary_conditions = boundary_conditions

    def calc_system_matrices(self):
        """Calculate and return system matrices in a dict."""
        which = 0
        space = None
        for form in self.matrix_forms.values():
            if isinstance(form, Forms.MatrixForm):
                which |= _BITS[form.matrix]
                space = space or form.function_space
        system = AssembledSystem(space, which) if which else None
        system_matrices = dict()
        for matname, form in self.matrix_forms.items():
            if isinstance(form, Forms.NullForm) or form is None:
                mat = None
            else:
                mat = DeviceMatrix(system, form.matrix)
                self.boundary_conditions.apply_essential(mat)
            system_matrices[matname] = mat
        return system_matrices
