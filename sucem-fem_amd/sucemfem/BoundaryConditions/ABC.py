"""First-order absorbing boundary condition (BoundaryConditions/ABC.py:26-58).

Bilinear form <n x v, n x u> over the whole exterior, assembled as S0 by
libsucem_hip (only the tangential traces of the face DOFs contribute).
"""
from .. import Forms
from .base_boundary_condition import BoundaryCondition


class ABCBoundaryCondition(BoundaryCondition):
    def get_bilinear_form(self, test_function=None, trial_function=None):
        return Forms.MatrixForm('S0', self.function_space)
