"""Field reconstruction (sucemfem/PostProcessing/PostProcessing.py:27-60).

Point location and basis evaluation run on the device; a point on a face shared by
several cells uses the lowest-index containing cell.
"""
import numpy as N

from ..Sources.point_source import basis_at_points

__all__ = ['Reconstruct']


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
