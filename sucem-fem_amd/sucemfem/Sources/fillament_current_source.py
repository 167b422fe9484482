"""Line-current filament source (Sources/fillament_current_source.py:27-82).

A straight filament carrying the current I from endpoint a to endpoint b is integrated
with the point rule the reference uses: n equally spaced points on the segment, both
endpoints included (the midpoint alone when n = 1), each carrying the dipole moment
I (b - a) / n.  The filament is that list of point dipoles and nothing else:
``get_point_sources`` hands it to ``CurrentSources`` / ``DrivenProblemABC.get_RHS``
(one ``sucem_rhs_point_sources`` launch for every source of a problem), and
``get_contribution`` sums the same points into a device vector with that kernel and
returns the sums on the DOFs of the cells the points lie in.
"""
from __future__ import division

import numpy as np

from .. import _hip
from .current_source import CurrentSource
from .point_source import basis_at_points


def filament_points(endpoints, current, n):
    """Quadrature points (n, 3) of the segment endpoints[0] -> endpoints[1] and the dipole
    moment (3,) each of them carries."""
    a, b = np.asarray(endpoints, np.float64).reshape(2, 3)
    s = np.array([0.5]) if n == 1 else np.linspace(0.0, 1.0, int(n))
    return a + np.outer(s, b - a), (b - a) * (current / float(n))


class FillamentCurrentSource(CurrentSource):
    no_integration_points = 100
    source_endpoints = None
    value = None

    def set_no_integration_points(self, no_integration_points):
        self.no_integration_points = int(no_integration_points)

    def set_source_endpoints(self, source_endpoints):
        """2x3 array: the current flows from the first point towards the second."""
        self.source_endpoints = np.asarray(source_endpoints, np.float64).reshape(2, 3)

    def set_value(self, value):
        """Filament current I."""
        self.value = value

    def get_point_sources(self):
        pts, moment = filament_points(self.source_endpoints, self.value,
                                      self.no_integration_points)
        return pts, np.tile(np.asarray(moment, np.complex128), (len(pts), 1))

    def get_contribution(self):
        """(dofnos, contributions): sum over the points of N_k(x) . moment, on the DOFs of
        the cells holding the points (first-seen order)."""
        V = self.function_space
        ctx = V.ctx
        pts, moments = self.get_point_sources()
        pts = np.ascontiguousarray(pts)
        b = ctx.tensor(V.dim())
        _hip.call("sucem_rhs_point_sources", ctx.handle, V.handle.h, len(pts), _hip.ptr(pts),
                  _hip.ptr(np.ascontiguousarray(moments)), None, 1, 1.0, 0.0, _hip.ptr(b))
        _, cell_dofs = basis_at_points(V, pts)
        flat = cell_dofs.reshape(-1)
        _, first = np.unique(flat, return_index=True)
        dofnos = flat[np.sort(first)]
        values = b.cpu().numpy()[dofnos]
        if not np.iscomplexobj(self.value):
            values = values.real.copy()
        return dofnos.astype(np.uint64), values
