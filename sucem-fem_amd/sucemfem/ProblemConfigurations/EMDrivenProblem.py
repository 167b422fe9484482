"""Driven problem with a first-order ABC (ProblemConfigurations/EMDrivenProblem.py:33-92).

A = S - k0^2 M + j k0 S_0 with k0 = 2 pi f / c0, and b = -j k0 Z0 sum N(x_s).p.
The matrices stay on the device: ``get_LHS_matrix`` exports the combination as
a scipy CSR matrix (for API compatibility) that carries the device operator, so
``Utilities.LinalgSolvers`` solves it with the fused on-the-fly SpMV instead of
uploading the exported values again.
"""
from __future__ import division

import numpy as N
import scipy.sparse

from .. import Forms
from .. import _hip
from ..Consts import c0, Z0
from .EMProblem import EMProblem


class CombineForms(Forms.CombineGalerkinForms):
    def get_forms(self):
        m = self.interior_forms.get_mass_form()
        s = self.interior_forms.get_stiffness_form()
        ABC_form = self.boundary_conditions.get_bilinear_form()
        return dict(M=m, S=s, S_0=ABC_form)


class DrivenLHS(scipy.sparse.csr_matrix):
    """The exported complex LHS; ``device_operator`` is (assembled system, k0)."""
    device_operator = None


class DrivenProblemABC(EMProblem):
    """Driven problem A = S - k0**2*M + 1j*k0*S_0 (lossless, frequency-independent materials)."""

    FormCombiner = CombineForms

    def set_sources(self, sources):
        self.sources = sources

    def set_frequency(self, frequency):
        """Set simulation frequency in Hz."""
        self.frequency = frequency

    def get_k0(self):
        return 2 * N.pi * self.frequency / c0

    def get_system(self):
        """The device-resident assembled system (S, M and S_0 share it)."""
        return self.system_matrices['S'].system

    def pec_diagonal(self, k0):
        """Diagonal of the PEC rows of A, or None when no essential condition was applied.

        DOLFIN's ``DirichletBC.apply`` (reached through ``calc_system_matrices``,
        SystemMatrices.py:46-58) zeroes each constrained row of S, M and S_0 separately and
        puts 1 on its diagonal, so a PEC row of S - k0^2 M + j k0 S_0 is
        (1 - k0^2 + j k0) e_i (terms of matrices left unconstrained drop out)."""
        mats = self.system_matrices
        on = dict((k, mats.get(k) is not None and mats[k].pec_constrained)
                  for k in ('S', 'M', 'S_0'))
        if not any(on.values()):
            return None
        return complex(1.0 if on['S'] else 0.0, 0.0) - (k0 ** 2 if on['M'] else 0.0) \
            + (1j * k0 if on['S_0'] else 0.0)

    def get_LHS_matrix(self):
        k0 = self.get_k0()
        system = self.get_system()
        indptr, indices = self.function_space.pattern()
        vals = system.lhs_values(k0)
        if self.system_matrices.get('S_0') is None:
            vals = vals.real.astype(N.complex128)
        d = self.pec_diagonal(k0)
        if d is not None:
            mask = self.function_space.pec_mask()
            rows = N.repeat(N.arange(len(indptr) - 1), N.diff(indptr))
            kill = mask[rows]
            vals[kill] = 0
            vals[kill & (indices == rows)] = d
        A = DrivenLHS((vals, indices.astype(N.int32), indptr.astype(N.int32)),
                      shape=(self.get_global_dimension(),) * 2)
        # SciPy's sparse sum in the reference drops exact zeros
        A.eliminate_zeros()
        A.device_operator = (system, k0, self.system_matrices.get('S_0') is not None, d)
        return A

    def get_RHS_device(self, nrhs=1):
        """Right-hand side as a complex128 device tensor (computed by the device kernel)."""
        k0 = self.get_k0()
        self.sources.set_function_space(self.function_space)
        self.sources.init_sources()
        pos, val = self.sources.get_point_sources()
        ctx = self.function_space.ctx
        b = ctx.tensor(self.get_global_dimension())
        pos = N.ascontiguousarray(pos, N.float64)
        val = N.ascontiguousarray(val, N.complex128)
        _hip.call("sucem_rhs_point_sources", ctx.handle, self.function_space.handle.h, len(pos),
                  _hip.ptr(pos), _hip.ptr(val), None, 1, 0.0, -k0 * Z0, _hip.ptr(b))
        return b

    def get_RHS(self):
        return self.get_RHS_device().cpu().numpy()

    def _get_RHS_contributions(self):
        self.sources.set_function_space(self.function_space)
        self.sources.init_sources()
        return self.sources.get_source_contributions()
