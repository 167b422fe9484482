"""Field reconstruction and the EM functional (sucemfem/PostProcessing/PostProcessing.py:27-172).

Point location and basis evaluation run on the device; a point on a face shared by
several cells uses the lowest-index containing cell.  The functional is evaluated with
the device operator (S and T assembled with per-cell material weights) and a device dot
product.
"""
import ctypes

import numpy as N

from .. import _hip
from ..Sources.point_source import basis_at_points
from ._device import device_dofs

__all__ = ['Reconstruct', 'CalcEMFunctional']


class Reconstruct(object):
    """Reconstruct field values E(x) = sum_i x_i N_i(x), complex where the DOFs are."""

    def __init__(self, function_space):
        self.function_space = function_space

    def set_dof_values(self, x):
        self.dof_values = N.asarray(x)

    def reconstruct_points(self, points):
        """Field at ``points`` (n x 3) -> n x 3 array."""
        vals, dofs = basis_at_points(self.function_space, N.asarray(points, N.float64))
        coeffs = self.dof_values[dofs]                    # (n, ndof_cell)
        out = N.einsum("pi,pid->pd", coeffs, vals)
        return out if N.iscomplexobj(out) else out.astype(N.float64)


def _per_cell(value, nc, name):
    """Per-cell values of a material function: None -> 1, a scalar, an array of one value
    per cell, or an entry of ``MaterialFunctionFactory.get_material_functions`` (region
    ids, values, region of each cell)."""
    if value is None:
        return N.ones(nc)
    if isinstance(value, tuple) and len(value) == 3:
        ids, vals, region = (N.asarray(v) for v in value)
        lookup = dict(zip(ids.tolist(), vals.tolist()))
        reg = N.zeros(nc, N.int64) if region.ndim == 0 or region.size == 0 else region
        return N.array([lookup[int(r)] for r in reg], N.float64)
    arr = N.asarray(value, N.float64)
    if arr.ndim == 0:
        return N.full(nc, float(arr))
    if arr.shape != (nc,):
        raise ValueError("%s: expected a scalar or %d per-cell values" % (name, nc))
    return arr


class CalcEMFunctional(object):
    """<curl E / mu_r, curl g> - k0^2 <eps_r E, g>, bilinear (no conjugation), over the
    whole mesh or over the cells of one domain mark (``PostProcessing.py:62-172``).

    Materials default to free space.  ``set_mur_function`` takes effect whether or not an
    eps_r function is set (the reference tests ``epsr_function`` there,
    ``PostProcessing.py:126-132``).  Integration is exact for cellwise-constant materials,
    so ``set_quadrature_degree`` is recorded and changes nothing.  The integral is
    g^T (S_w - k0^2 T_w) E with S_w, T_w assembled on a private copy of the mesh whose
    per-cell weights are 1/mu_r and eps_r on the marked cells and zero elsewhere.
    """

    def __init__(self, function_space, testing_space=None):
        if testing_space is not None and testing_space is not function_space:
            raise NotImplementedError("the testing space must be the solution space")
        self.function_space = function_space
        self.testing_space = function_space
        self.cell_domains = None
        self.cell_mark_value = None
        self.epsr_function = None
        self.mur_function = None
        self.quadrature_degree = None
        self._system = None
        self.dirty = True

    def set_k0(self, k0):
        self.k0 = k0

    def set_cell_domains(self, cell_domains, cell_mark_value):
        """Restrict the integral to the cells with ``cell_domains[c] == cell_mark_value``."""
        self.cell_domains = cell_domains
        self.cell_mark_value = cell_mark_value
        self.dirty = True

    def set_epsr_function(self, epsr_function):
        self.epsr_function = epsr_function
        self.dirty = True

    def set_mur_function(self, mur_function):
        self.mur_function = mur_function
        self.dirty = True

    def set_quadrature_degree(self, quadrature_degree):
        self.quadrature_degree = quadrature_degree

    def set_E_dofs(self, E_dofs):
        self._E = device_dofs(self.function_space, E_dofs)

    def set_g_dofs(self, g_dofs):
        self._g = device_dofs(self.testing_space, g_dofs)

    def _weighted_system(self):
        if self._system is not None and not self.dirty:
            return self._system
        from ..Mesh import Mesh, FunctionSpace
        from ..SystemMatrices import AssembledSystem
        V = self.function_space
        mesh = V.mesh()
        nc = mesh.num_cells()
        eps = _per_cell(self.epsr_function, nc, "eps_r")
        mu = _per_cell(self.mur_function, nc, "mu_r")
        if N.any(mu == 0):
            raise ValueError("mu_r must be nonzero")
        inv_mu = 1.0 / mu
        if self.cell_domains is not None:
            marks = N.asarray(self.cell_domains.array() if hasattr(self.cell_domains, "array")
                              else self.cell_domains)
            if marks.shape != (nc,):
                raise ValueError("cell_domains must hold one value per cell")
            inside = marks == self.cell_mark_value
            eps = N.where(inside, eps, 0.0)
            inv_mu = N.where(inside, inv_mu, 0.0)
        copy = Mesh(mesh.coordinates(), mesh.cells())
        mh = copy.device_handle(V.ctx)
        eps = N.ascontiguousarray(eps, N.float64)
        inv_mu = N.ascontiguousarray(inv_mu, N.float64)
        _hip.call("sucem_mesh_set_cell_materials", mh.h, _hip.ptr(eps), _hip.ptr(inv_mu))
        space = FunctionSpace(copy, "Nedelec 1st kind H(curl)", V.order, ctx=V.ctx)
        self._system = (copy, space, AssembledSystem(space, _hip.MAT_S | _hip.MAT_T))
        self.dirty = False
        return self._system

    def calc_functional(self):
        """The functional for the current E, g and k0 (complex)."""
        import torch
        _, space, system = self._weighted_system()
        y = torch.empty_like(self._E)
        k0 = float(self.k0)
        _hip.call("sucem_system_apply_coef", space.ctx.handle, system.handle.h, 1.0, -k0 * k0,
                  0.0, 64, 0, 1, _hip.ptr(self._E), _hip.ptr(y))
        return complex(torch.dot(self._g, y).item())
