"""Weak-form descriptors (sucemfem/Forms.py:25-116).

The reference builds UFL forms that DOLFIN integrates; here a form is a
descriptor naming the matrix libsucem_hip assembles for it:
``eps_r*inner(v,u)*dx`` -> T, ``dot(curl v, curl u)/mu_r*dx`` -> S,
``inner(cross(n,v), cross(n,u))*ds`` -> S0.
"""


class NullForm(object):
    """Zero form; adding it to another form gives that form (Forms.py:25-31)."""
    _integrals = []
    matrix = None

    def __add__(self, other):
        return other

    def __radd__(self, other):
        return other


class MatrixForm(object):
    """Descriptor of one assembled bilinear form."""

    def __init__(self, matrix, function_space=None, material_functions=None):
        self.matrix = matrix            # 'T', 'S' or 'S0'
        self.function_space = function_space
        self.material_functions = material_functions

    def __add__(self, other):
        if isinstance(other, NullForm):
            return self
        if isinstance(other, MatrixForm) and other.matrix == self.matrix:
            return self
        raise NotImplementedError("sums of different bilinear forms are not supported")

    def __repr__(self):
        return "MatrixForm(%s)" % self.matrix


class GalerkinInteriorForms(object):
    def set_material_functions(self, material_functions):
        self.material_functions = material_functions

    def set_function_space(self, function_space):
        self.function_space = function_space

    def get_mass_form(self):
        raise NotImplementedError()

    def get_stiffness_form(self):
        raise NotImplementedError()


class EMGalerkinInteriorForms(GalerkinInteriorForms):
    """Mass ``eps_r v.u`` and stiffness ``curl v . curl u / mu_r`` (Forms.py:65-85)."""

    def get_mass_form(self):
        return MatrixForm('T', self.function_space, self.material_functions)

    def get_stiffness_form(self):
        return MatrixForm('S', self.function_space, self.material_functions)


class CombineGalerkinForms(object):
    def set_interior_forms(self, interior_forms):
        self.interior_forms = interior_forms

    def set_boundary_conditions(self, boundary_conditions):
        self.boundary_conditions = boundary_conditions

    def get_combined_forms(self):
        raise NotImplementedError("Use a concrete class")
