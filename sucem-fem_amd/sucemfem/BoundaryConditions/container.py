"""Collection of boundary conditions keyed by region number (BoundaryConditions/container.py:25-68)."""
from __future__ import division

from .. import Forms


class BoundaryConditions(object):
    def __init__(self):
        self.boundary_conditions = {}

    def add_boundary_condition(self, boundary_condition):
        bc_num = boundary_condition.region_number
        assert bc_num not in self.boundary_conditions
        self.boundary_conditions[bc_num] = boundary_condition

    def apply_essential(self, A, b=None):
        """Apply the essential boundary conditions to system matrix A (and optional RHS b)."""
        for bc_num, bc in self.boundary_conditions.items():
            apply_fn = bc.get_essential_application_func()
            if b is not None:
                apply_fn(A, b)
            else:
                apply_fn(A)

    def get_linear_form(self):
        lin_form = Forms.NullForm()
        for bc_num, bc in self.boundary_conditions.items():
            lin_form = lin_form + bc.get_linear_form()
        return lin_form

    def get_bilinear_form(self):
        bilin_form = Forms.NullForm()
        for bc_num, bc in self.boundary_conditions.items():
            bilin_form = bilin_form + bc.get_bilinear_form()
        return bilin_form

    def set_function_space(self, function_space):
        self.function_space = function_space
        for bc in self.boundary_conditions.values():
            bc.set_function_space(function_space)

    def has_essential(self):
        from .essential import EssentialBoundaryCondition
        return any(isinstance(bc, EssentialBoundaryCondition)
                   for bc in self.boundary_conditions.values())
