"""Essential (PEC) boundary conditions (BoundaryConditions/essential.py:24-99).

DOLFIN's topological ``DirichletBC.apply`` constrains every DOF on the closure of
the marked facets: it zeroes the matrix row and puts 1 on the diagonal.  Here
the constraint is recorded on the assembled device matrices; solvers eliminate
the constrained DOFs (restrict to the free DOFs) and exports reproduce the
zero-row / unit-diagonal form.
"""
from __future__ import division

import numpy as np

from .base_boundary_condition import BoundaryCondition


class EssentialBoundaryCondition(BoundaryCondition):
    def init_with_meshfunction(self, mesh_function, region_number):
        self.set_region_number(region_number)
        self.set_mesh_function(mesh_function)

    def set_PEC_expression(self):
        self.set_boundary_value_expression(('0.0', '0.0', '0.0'))

    def get_essential_application_func(self, function_space=None):
        if function_space is not None:
            self.set_function_space(function_space)
        if self.boundary_value_expression is None:
            self.set_PEC_expression()
        if any(v not in ('0.0', 0, 0.0) for v in self.boundary_value_expression):
            raise NotImplementedError("only homogeneous (PEC) essential conditions are supported")

        def apply(A, b=None):
            self._mark_space()
            if hasattr(A, "constrain_pec"):
                A.constrain_pec()
            if b is not None:
                mask = self.function_space.pec_mask()
                b[mask] = 0
        return apply


    def _mark_space(self):
        """Set the space's PEC DOFs from this condition's facet marks (once per space);
        several essential conditions on one space add up."""
        V = self.function_space
        if getattr(V, "_pec_marked_by", None) is None:
            V._pec_marked_by = []
        if self in V._pec_marked_by:
            return
        union = len(V._pec_marked_by) > 0
        mf = self.mesh_function
        if mf is None:
            V.mark_pec(None, union=union)
        else:
            if getattr(mf, "dim", 2) != 2:
                raise ValueError("essential conditions need a facet (dim 2) mesh function")
            V.mark_pec(np.asarray(mf.array()) == self.region_number, union=union)
        V._pec_marked_by.append(self)


class PECWallsBoundaryCondition(EssentialBoundaryCondition):
    """PEC on every boundary facet (essential.py:86-99: DomainBoundary marked 999)."""

    def init_with_mesh(self, mesh):
        self.init_with_meshfunction(None, 999)
