"""Line-current filament as a row of point sources (Sources/fillament_current_source.py:27-82)."""
from __future__ import division

import collections

import numpy as np

from .current_source import CurrentSource
from .point_source import basis_at_points


def unit_vector(v):
    v = np.asarray(v, np.float64)
    return v / np.sqrt(np.sum(v ** 2))


def vector_length(v):
    return np.sqrt(np.sum(np.asarray(v, np.float64) ** 2))


class FillamentCurrentSource(CurrentSource):
    no_integration_points = 100

    def __init__(self, *names, **kwargs):
        self._dirty = True

    def set_no_integration_points(self, no_integration_points):
        self.no_integration_points = no_integration_points

    def set_source_endpoints(self, source_endpoints):
        """2x3 array: current flows from the first point towards the second."""
        self.source_endpoints = np.asarray(source_endpoints, np.float64)
        self._dirty = True

    def set_value(self, value):
        self.value = value
        self._dirty = True

    def _update(self):
        self.source_start, self.source_end = self.source_endpoints
        self.source_delta = self.source_end - self.source_start
        self.vector_value = unit_vector(self.source_delta) * self.value

    def get_point_sources(self):
        self._update()
        source_len = vector_length(self.source_delta)
        no_pts = self.no_integration_points
        if no_pts == 1:
            pts = (self.source_start + self.source_delta * 0.5)[None, :]
        else:
            pts = self.source_start + self.source_delta * np.linspace(0, 1, no_pts)[:, np.newaxis]
        point_magnitude = self.vector_value * source_len / no_pts
        return pts, np.repeat(np.asarray(point_magnitude, np.complex128)[None, :], len(pts), axis=0)

    def get_contribution(self):
        pts, mags = self.get_point_sources()
        vals, dofs = basis_at_points(self.function_space, pts)
        mag = np.real_if_close(mags[0])
        contribs = collections.OrderedDict()
        for k in range(len(pts)):
            for dn, v in zip(dofs[k], vals[k] @ mag):
                contribs[dn] = contribs.get(dn, 0.) + v
        dofnos = np.array(list(contribs.keys()), dtype=np.uint64)
        rhs_contribs = np.array(list(contribs.values()))
        return dofnos, rhs_contribs
