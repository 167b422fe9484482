"""Variational near-to-far-field transform (sucemfem/PostProcessing/variational_ntff.py:31-131).

The H contribution is the free-space functional <curl E, curl g> - k0^2 <E, g> over the
cells that touch the boundary through an edge, g the boundary interpolant of the
testing field (j/k0) r x (a x r) exp(j k0 r.x); the E contribution is the surface L
potential.  Both run on the device (``sucem_ntff`` with variational = 1).
"""
from __future__ import division

import numpy as np

from . import surface_ntff

__all__ = ['NTFF', 'TransformTestingExpression']


class _Component(object):
    def __init__(self, owner, part):
        self.owner, self.part = owner, part

    def eval(self, values, x):
        v = self.owner(np.asarray(x, np.float64)[None, :])[0]
        values[:] = v.real if self.part == 0 else v.imag


class TransformTestingExpression(object):
    """(j/k0) rhat x (ahat x rhat) exp(j k0 rhat.x) with the reference's
    ``set_parms``/``get_expression`` interface (``variational_ntff.py:31-61``); the
    expressions expose ``eval(values, x)``."""

    def set_parms(self, rhat, ahat, k0):
        self.rhat = np.asarray(rhat, np.float64)
        self.cf = np.cross(self.rhat, np.cross(ahat, self.rhat)) / k0
        self.k0 = k0

    def __call__(self, pts):
        return 1j * self.cf[None, :] * np.exp(1j * self.k0 * (pts @ self.rhat))[:, None]

    def get_expression(self):
        return _Component(self, 0), _Component(self, 1)


class NTFF(surface_ntff.NTFF):
    _variational = 1

    def __init__(self, function_space, testing_space=None):
        if testing_space is not None and testing_space is not function_space:
            raise NotImplementedError("the testing space is the solution space")
        super(NTFF, self).__init__(function_space)

    def set_k0(self, k0):
        self.k0 = k0
