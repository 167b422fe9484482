"""Current-source container (sucemfem/Sources/current_source.py:27-66)."""
from __future__ import division

import collections

import numpy as N


class CurrentSources(object):
    def __init__(self):
        self.sources = []

    def set_function_space(self, function_space):
        self.function_space = function_space

    def add_source(self, source):
        self.sources.append(source)

    def init_sources(self):
        for src in self.sources:
            src.set_function_space(self.function_space)

    def get_source_contributions(self):
        """(dofnos, contributions) summed over all sources, DOFs in first-seen order."""
        contribs = collections.OrderedDict()
        for src in self.sources:
            dofnos, values = src.get_contribution()
            for dn, v in zip(dofnos, values):
                contribs[dn] = contribs.get(dn, 0.) + v
        dofnos = N.array(list(contribs.keys()), dtype=N.uint64)
        rhs_contribs = N.array(list(contribs.values()))
        return dofnos, rhs_contribs

    def get_point_sources(self):
        """All sources as point dipoles: (positions (n,3), complex moments (n,3))."""
        pos, val = [], []
        for src in self.sources:
            p, v = src.get_point_sources()
            pos.append(N.asarray(p, N.float64).reshape(-1, 3))
            val.append(N.asarray(v, N.complex128).reshape(-1, 3))
        if not pos:
            return N.zeros((0, 3)), N.zeros((0, 3), N.complex128)
        return N.concatenate(pos), N.concatenate(val)


class CurrentSource(object):
    """Abstract base class for current sources."""

    def set_function_space(self, function_space):
        self.function_space = function_space

    def get_contribution(self):
        raise NotImplementedError('User subclass should implement get_contribution()')

    def get_point_sources(self):
        raise NotImplementedError('User subclass should implement get_point_sources()')
