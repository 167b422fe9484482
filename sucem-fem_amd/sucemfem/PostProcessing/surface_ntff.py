"""Surface near-to-far-field transform (sucemfem/PostProcessing/surface_ntff.py:125-155).

Equivalent currents J = n x H, M = -n x E on the mesh boundary, integrated on the
device against exp(j k0 r'.r_hat) (``sucem_ntff``); far fields are normalised to
radius 1.
"""
from __future__ import division

import numpy as np

from .. import _hip
from ..Consts import c0
from ._device import device_dofs

__all__ = ['NTFF']


def calc_points(space, x, k0, theta_deg, phi_deg, variational):
    theta = np.ascontiguousarray(np.deg2rad(np.atleast_1d(theta_deg)), np.float64)
    phi = np.ascontiguousarray(np.deg2rad(np.atleast_1d(phi_deg)), np.float64)
    if theta.shape != phi.shape:
        raise ValueError("theta and phi must have the same length")
    out = np.zeros((len(theta), 2), np.complex128)
    _hip.call("sucem_ntff", space.ctx.handle, space.handle.h, _hip.ptr(x), float(k0),
              int(variational), len(theta), _hip.ptr(theta), _hip.ptr(phi), _hip.ptr(out))
    return out


class NTFF(object):
    _variational = 0

    def __init__(self, function_space):
        self.function_space = function_space

    def set_dofs(self, dofs):
        self._x = device_dofs(self.function_space, dofs)

    def set_frequency(self, frequency):
        self.frequency = frequency
        self.k0 = self.frequency * 2 * np.pi / c0

    def calc_pt(self, theta_deg, phi_deg):
        """(E_theta, E_phi) at one direction, degrees."""
        e = calc_points(self.function_space, self._x, self.k0, theta_deg, phi_deg,
                        self._variational)[0]
        return (e[0], e[1])

    def calc_pts(self, theta_deg, phi_deg):
        """All directions in one device pass: complex (n, 2)."""
        return calc_points(self.function_space, self._x, self.k0, theta_deg, phi_deg,
                           self._variational)
