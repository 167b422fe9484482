"""Power flux through the mesh boundary (sucemfem/PostProcessing/power_flux.py:11-133).

Both classes keep the reference's interface (``set_dofs``, ``set_k0``,
``calc_flux``); the integrals run on the device through ``sucem_surface_flux`` and
``sucem_variational_flux`` on the solution vector resident in HBM.
"""
from __future__ import division

import ctypes

import numpy as np

from .. import _hip
from ._device import device_dofs

__all__ = ['SurfaceFlux', 'VariationalSurfaceFlux']


class SurfaceFlux(object):
    """Real part of the Poynting flux through the boundary from a surface integral
    (``power_flux.py:11-63``); peak values, so the RMS power is half of it."""

    def __init__(self, function_space):
        self.function_space = function_space
        self.mur_function = None

    def set_dofs(self, dofs):
        self._x = device_dofs(self.function_space, dofs)

    def set_k0(self, k0):
        self.k0 = k0

    def set_mur_function(self, mur_function):
        """Any non-None value uses the mesh's per-cell mu_r (set with the materials)."""
        self.mur_function = mur_function

    def calc_flux(self):
        V = self.function_space
        out = ctypes.c_double()
        _hip.call("sucem_surface_flux", V.ctx.handle, V.handle.h, _hip.ptr(self._x),
                  float(self.k0), int(self.mur_function is not None), ctypes.byref(out))
        return out.value


class VariationalSurfaceFlux(object):
    """Flux from the free-space variational functional over the cells that touch the
    boundary through an edge (``power_flux.py:65-133``); complex, as the reference's."""

    def __init__(self, function_space):
        self.function_space = function_space

    def set_dofs(self, dofs):
        self._x = device_dofs(self.function_space, dofs)
        self.dofs = dofs

    def set_k0(self, k0):
        self.k0 = k0

    def set_epsr_function(self, epsr_function):
        raise NotImplementedError("the variational flux is evaluated in free space "
                                  "(power_flux.py:105-118 never applies eps_r)")

    def set_mur_function(self, mur_function):
        raise NotImplementedError("the variational flux is evaluated in free space")

    def calc_flux(self):
        V = self.function_space
        out = np.zeros(2)
        _hip.call("sucem_variational_flux", V.ctx.handle, V.handle.h, _hip.ptr(self._x),
                  float(self.k0), _hip.ptr(out))
        return complex(out[0], out[1])
