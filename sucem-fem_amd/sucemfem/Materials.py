"""Material regions -> per-cell eps_r and 1/mu_r (sucemfem/Materials.py:30-150).

The region table is resolved per cell on the device by
``sucem_mesh_set_materials``; an unknown region number raises ``ValueError``
as at Materials.py:147-148.
"""
from __future__ import division

import numpy as np

from . import Consts
from . import _hip

mu0 = Consts.mu0
eps0 = Consts.eps0


class MaterialProperties(object):
    allowed_properties_defaults = {'eps_r': 1.0, 'mu_r': 1.0}

    def __init__(self):
        for prop, val in self.allowed_properties_defaults.items():
            setattr(self, prop, val)

    def init_values(self, **kwargs):
        for prop, val in kwargs.items():
            if prop not in self.allowed_properties_defaults:
                raise ValueError('Unknown material property')
            setattr(self, prop, val)

    def set_eps_r(self, eps_r):
        self.eps_r = eps_r

    def set_mu_r(self, mu_r):
        self.mu_r = mu_r

    def get_eps_r(self):
        return self.eps_r

    def get_mu_r(self):
        return self.mu_r

    def get_mu_r_inv(self):
        return 1 / self.get_mu_r()

    def get_eps(self):
        return eps0 * self.eps_r

    def get_mu(self):
        return mu0 * self.mu_r


class MaterialPropertiesFactory(object):
    """Region number -> MaterialProperties (Materials.py:64-95); None -> region 0 free space."""

    def __init__(self, material_regions):
        self.material_properties = {}
        if material_regions is None:
            self.material_properties[0] = MaterialProperties()
        else:
            for region_no, property_values in material_regions.items():
                matprop = MaterialProperties()
                matprop.init_values(**property_values)
                self.material_properties[region_no] = matprop

    def get_material_properties(self):
        return self.material_properties


class MaterialFunctions(dict):
    """Per-cell material data living on the device mesh ({'eps_r': ..., 'mu_r': ...})."""


class MaterialFunctionFactory(object):
    """Per-cell material functions on the device (Materials.py:118-150)."""

    def __init__(self, region_material_properties, region_meshfunction, mesh):
        self.region_material_properties = region_material_properties
        self.region_meshfunction = region_meshfunction
        self.mesh = mesh

    def region_table(self):
        ids = np.array(sorted(self.region_material_properties), np.int64)
        eps = np.array([self.region_material_properties[i].get_eps_r() for i in ids], np.float64)
        mu = np.array([self.region_material_properties[i].get_mu_r() for i in ids], np.float64)
        return ids, eps, mu

    def get_material_functions(self, *property_names, **kw):
        """Resolve regions per cell on the device; returns a dict keyed by property name."""
        for p in property_names:
            if p not in MaterialProperties.allowed_properties_defaults:
                raise ValueError('Unknown material property')
        ctx = kw.get("ctx")
        mh = self.mesh.device_handle(ctx)
        ids, eps, mu = self.region_table()
        region = None
        if self.region_meshfunction is not None:
            arr = self.region_meshfunction.array() if hasattr(self.region_meshfunction, "array") \
                else self.region_meshfunction
            region = np.ascontiguousarray(arr, np.int64)
            if len(region) != self.mesh.num_cells():
                raise ValueError('region meshfunction does not match the mesh')
        _hip.call("sucem_mesh_set_materials", mh.h, _hip.ptr(region), len(ids), _hip.ptr(ids),
                  _hip.ptr(eps), _hip.ptr(mu))
        out = MaterialFunctions()
        for p in property_names:
            out[p] = (ids, eps if p == 'eps_r' else mu, region)
        return out
