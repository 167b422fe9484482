"""Filament-source parameter helper (sucemfem/Sources/fillament_source.py)."""
from __future__ import division

import numpy as np

from .fillament_current_source import FillamentCurrentSource


class FillamentSource(object):
    def __init__(self, function_space):
        self.function_space = function_space

    def set_source_parameters(self, source_parameters):
        self.source_parameters = source_parameters
        self.I = source_parameters['I']
        self.endpoints = np.asarray(source_parameters['endpoints'], np.float64)
        delta = self.endpoints[1] - self.endpoints[0]
        self.length = np.sqrt(np.sum(delta ** 2))
        self.direction = delta / self.length

    def get_current_source(self):
        cs = FillamentCurrentSource()
        cs.set_function_space(self.function_space)
        cs.set_source_endpoints(self.endpoints)
        cs.set_value(self.I)
        return cs
