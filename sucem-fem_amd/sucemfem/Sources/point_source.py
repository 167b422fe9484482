"""Point (infinitesimal dipole) current source (sucemfem/Sources/point_source.py:28-94).

Cell location and basis evaluation run on the device (``sucem_locate_points``,
``sucem_basis_at_points``).  A point on a face, edge or vertex shared by several
cells is assigned to the lowest-index containing cell; a point outside the mesh
raises ``AssertionError`` as at point_source.py:79.
"""
from __future__ import division

import numpy as N

from .. import _hip
from .current_source import CurrentSource


class PointCurrentSource(CurrentSource):
    def set_position(self, position):
        self.position = N.array(position, dtype=N.float64)

    def set_value(self, value):
        iscomplex = N.any(N.iscomplex(value))
        dtype = N.complex128 if iscomplex else N.float64
        self.value = N.array(value, dtype=dtype)

    def get_contribution(self):
        return calc_pointsource_contrib(self.function_space, self.position, self.value)

    def get_point_sources(self):
        return self.position[None, :], N.asarray(self.value, N.complex128)[None, :]


def basis_at_points(V, points):
    """Physical basis values (n, ndof_cell, 3) and cell DOFs (n, ndof_cell) at points."""
    pts = _hip.f64(points, (-1, 3))
    nloc = V.dofmap().max_cell_dimension()
    vals = N.empty((len(pts), nloc, 3), N.float64)
    dofs = N.empty((len(pts), nloc), N.int64)
    _hip.call("sucem_basis_at_points", V.ctx.handle, V.handle.h, len(pts), _hip.ptr(pts),
              _hip.ptr(vals), _hip.ptr(dofs))
    return vals, dofs


def calc_pointsource_contrib(V, source_coords, source_value):
    """(dofnos, sum_k N_k(x_s) . value) for the cell containing ``source_coords``."""
    vals, dofs = basis_at_points(V, N.asarray(source_coords, N.float64)[None, :])
    source_value = N.asarray(source_value)
    rhs_contribs = N.sum(vals[0] * source_value, axis=1)
    return dofs[0].astype(N.uint64), rhs_contribs


def containing_cells_midpoints(mesh, p, tol=1e-12):
    """Cells whose closure contains ``p`` and their midpoints (mesh centring helper)."""
    X = mesh.coordinates()[mesh.cells()]
    J = N.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)
    lam = N.linalg.solve(J, (p - X[:, 0])[:, :, None])[:, :, 0]
    lam = N.concatenate([1 - lam.sum(axis=1, keepdims=True), lam], axis=1)
    cand = N.flatnonzero(N.all(lam >= -tol, axis=1))
    if len(cand) == 0:
        raise AssertionError("point %r is outside the mesh" % (p,))
    return cand, X[cand].mean(axis=1)
