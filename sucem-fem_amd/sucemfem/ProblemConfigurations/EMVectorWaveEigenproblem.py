"""Resonant-cavity eigenproblem (ProblemConfigurations/EMVectorWaveEigenproblem.py:38-105).

``DefaultEigenSolver.solve_problem(nev)`` runs the device shift-invert Lanczos
(``sucem_eigs``) on the PEC-eliminated pencil and, like the reference, returns
``(eigs_w, eigs_v)`` with ``eigs_w = nu = 1/(k^2 - sigma)`` (callers recover k^2
as ``1/eigs_w + sigma``, examples/resonant_cavity/driver.py:66) and ``eigs_v``
of shape (nev, N).  ``solve_problem_k2`` returns k^2 directly.  The modes are
the nev lowest k^2 above sigma; the gradient null space (k^2 = 0) is excluded.
"""
import ctypes

import numpy as np

from .. import Forms
from .. import _hip
from .EMProblem import EMProblem


class CombineForms(Forms.CombineGalerkinForms):
    def get_forms(self):
        m = self.interior_forms.get_mass_form()
        s = self.interior_forms.get_stiffness_form()
        BC_bilin_forms = self.boundary_conditions.get_bilinear_form()
        return dict(M=m, S=s, BC=BC_bilin_forms)


class EigenProblem(EMProblem):
    FormCombiner = CombineForms


class DefaultEigenSolver(object):
    def __init__(self):
        self.sigma = 0.0
        self.options = {}
        self.stats = None

    def set_eigenproblem(self, eigenproblem):
        """Sets initialised instance of EigenProblem to solve."""
        self.eigenproblem = eigenproblem

    def set_sigma(self, sigma):
        """Spectrum shift (sigma) to apply to k^2."""
        self.sigma = sigma

    def set_options(self, **options):
        """ncv, max_restarts, inner_precision (64|32), inner_max_iter, tol, inner_tol, seed,
        block (Lanczos block size 1 or 2; 0, the default, picks 2 where paired inner solves
        run)."""
        self.options.update(options)

    def solve_problem_k2(self, nev, ncv=None, return_vectors=True):
        """(k^2 ascending, modes (nev, N) or None, relative residuals)."""
        ep = self.eigenproblem
        S = ep.system_matrices['S']
        system = S.system
        V = ep.function_space
        ctx = V.ctx
        o = _hip.EigOpts()
        _hip.load().sucem_eig_opts_default(ctypes.byref(o))
        if ncv is not None:
            o.ncv = int(ncv)
        for k, v in self.options.items():
            setattr(o, k, v)
        k2 = np.zeros(nev, np.float64)
        resid = np.zeros(nev, np.float64)
        vecs = ctx.tensor(nev * V.dim()) if return_vectors else None
        st = _hip.EigStats()
        _hip.call("sucem_eigs", ctx.handle, system.handle.h, float(self.sigma), int(nev),
                  ctypes.byref(o), _hip.ptr(k2), _hip.ptr(resid), _hip.ptr(vecs), ctypes.byref(st))
        self.stats = st.as_dict()
        modes = None if vecs is None else vecs.reshape(nev, V.dim()).cpu().numpy()
        return k2, modes, resid

    def solve_problem(self, nev, ncv=None):
        """Solve and return (eigs_w, eigs_v) as the reference does (eigs_w = 1/(k^2 - sigma))."""
        k2, modes, _ = self.solve_problem_k2(nev, ncv)
        eigs_w = (1.0 / (k2 - self.sigma)).astype(np.complex128)
        return eigs_w, modes
