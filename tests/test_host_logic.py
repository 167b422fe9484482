"""Host-side logic of the reference-shaped API (no GPU calls)."""
import os
import tempfile
import warnings

import numpy as np
import pytest
import scipy.sparse


def test_constants_follow_reference_evaluation_order():
    from sucemfem import Consts
    assert Consts.c0 == 299792458
    assert Consts.mu0 == np.pi * 4 * 1e-7
    assert Consts.Z0 == Consts.mu0 * Consts.c0
    assert abs(Consts.eps0 * Consts.mu0 * Consts.c0 ** 2 - 1) < 1e-15


def test_material_properties():
    from sucemfem.Materials import MaterialProperties, MaterialPropertiesFactory
    m = MaterialProperties()
    assert m.get_eps_r() == 1.0 and m.get_mu_r_inv() == 1.0
    with pytest.raises(ValueError):
        m.init_values(sigma=3)
    f = MaterialPropertiesFactory({100: {'eps_r': 2, 'mu_r': 4}})
    p = f.get_material_properties()[100]
    assert p.get_eps_r() == 2 and p.get_mu_r_inv() == 0.25
    assert list(MaterialPropertiesFactory(None).get_material_properties()) == [0]


def test_material_function_table():
    from sucemfem.Materials import MaterialPropertiesFactory, MaterialFunctionFactory
    props = MaterialPropertiesFactory({3: dict(eps_r=4.4), 1: dict(eps_r=2.2, mu_r=2)})
    fac = MaterialFunctionFactory(props.get_material_properties(), None, None)
    ids, eps, mu = fac.region_table()
    assert list(ids) == [1, 3] and list(eps) == [2.2, 4.4] and list(mu) == [2, 1]
    with pytest.raises(ValueError):
        fac.get_material_functions('sigma')


def test_forms_and_null_form():
    from sucemfem import Forms
    nf = Forms.NullForm()
    f = Forms.MatrixForm('S0')
    assert (nf + f) is f and (f + nf) is f
    assert (nf + nf) is not None
    with pytest.raises(NotImplementedError):
        Forms.GalerkinInteriorForms().get_mass_form()
    with pytest.raises(NotImplementedError):
        Forms.CombineGalerkinForms().get_combined_forms()


def test_boundary_condition_container():
    from sucemfem.BoundaryConditions import ABCBoundaryCondition, BoundaryConditions
    from sucemfem.BoundaryConditions import PECWallsBoundaryCondition
    bcs = BoundaryConditions()
    abc = ABCBoundaryCondition()
    abc.set_region_number(1)
    bcs.add_boundary_condition(abc)
    with pytest.raises(AssertionError):
        bcs.add_boundary_condition(abc)
    assert bcs.get_bilinear_form().matrix == 'S0'
    assert not bcs.has_essential()
    pec = PECWallsBoundaryCondition()
    pec.init_with_mesh(None)
    assert pec.region_number == 999
    bcs.add_boundary_condition(pec)
    assert bcs.has_essential()


def test_fillament_source_parameters():
    """test_fillament_source.py:94-104."""
    from sucemfem.Sources.fillament_source import FillamentSource
    from sucemfem.Sources.fillament_current_source import FillamentCurrentSource
    fs = FillamentSource(None)
    ep = np.array([[0, 0, 0.7], [0, 0, 0.2]])
    fs.set_source_parameters(dict(I=2, endpoints=ep))
    assert np.all(fs.direction == [0, 0, -1]) and abs(fs.length - 0.5) < 1e-15
    cs = fs.get_current_source()
    assert isinstance(cs, FillamentCurrentSource) and cs.value == 2
    assert np.all(cs.source_endpoints == ep)
    pts, mags = cs.get_point_sources()
    assert pts.shape == (100, 3) and np.allclose(mags.sum(axis=0).real, [0, 0, -2 * 0.5])


def test_point_source_values():
    from sucemfem.Sources.point_source import PointCurrentSource
    s = PointCurrentSource()
    s.set_value([1, 2, 3])
    assert s.value.dtype == np.float64
    s.set_value([1j, 0, 0])
    assert s.value.dtype == np.complex128


def test_calculate_residual_shapes():
    """test_LinalgSolvers.py:59-95: residual of the identity system with 1-D / column / row b."""
    from sucemfem.Utilities.LinalgSolvers import calculate_residual
    n = 50
    A = scipy.sparse.eye(n, n)
    rng = np.random.default_rng(0)
    for bshape, xshape in (((n,), (n,)), ((1, n), (n,)), ((n, 1), (n,)), ((n,), (n, 1)),
                           ((1, n), (n, 1)), ((n, 1), (n, 1))):
        b = rng.random(bshape)
        x = b.copy().reshape(xshape)
        assert calculate_residual(A, x, b) == 0


def test_matrix_io_roundtrip():
    """test_MatrixIO.py:39-84."""
    from sucemfem.Utilities import MatrixIO
    A = scipy.sparse.random(30, 30, density=0.1, random_state=1, format="csr")
    with tempfile.TemporaryDirectory() as d:
        assert MatrixIO.save_scipy_matrix_as_mat(d, "A", A)
        B = MatrixIO.load_scipy_matrix_from_mat(d, "A")
        assert abs(A - scipy.sparse.csr_matrix(B)).max() == 0
        assert MatrixIO.load_scipy_matrix_from_mat(d, "missing") is None


def test_solver_preconditioner_selection_without_gpu():
    from sucemfem.Utilities.LinalgSolvers import BiCGStabSolver, COCGSolver
    from sucemfem import _hip
    A = scipy.sparse.eye(4, format="csr")
    s = BiCGStabSolver(A, 'diagonal')
    assert s._precond == _hip.PRECOND_JACOBI
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        s = BiCGStabSolver(A, 'ilu')
        assert s._precond == _hip.PRECOND_JACOBI and w
        s = COCGSolver(A, 'unknown')
        assert s._precond == _hip.PRECOND_NONE
    assert COCGSolver(A, 'gradient')._precond == _hip.PRECOND_GRADIENT
    assert COCGSolver(A, 'auxiliary')._precond == _hip.PRECOND_AUXILIARY
    log = s.get_logging_data()
    assert log['id'][:3] == ['init', 'preconditioner::start', 'preconditioner::end']


def test_device_api_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from sucemfem.runtime import get_context
    with pytest.raises(RuntimeError):
        get_context()


def test_device_binary_searches_use_the_overflow_free_midpoint():
    """Row pointers of the 120 x 60 x 90-box order-2 pattern pass 2^30 (1,078,395,480
    nonzeros), so a 32-bit midpoint (lo + hi) / 2 wraps negative there.  Every CSR row
    search in the kernels goes through find_col, whose midpoint is lo + (hi - lo) / 2."""
    import glob
    import re
    root = os.path.join(os.path.dirname(__file__), "..", "sucem-fem_amd", "csrc")
    summed = re.compile(r"\(\s*(lo|low|a)\s*\+\s*(hi|high|b)\s*\)\s*(>>\s*1|/\s*2)")
    offenders = []
    for path in sorted(glob.glob(os.path.join(root, "*.hip")) + glob.glob(os.path.join(root, "*.hpp"))):
        with open(path) as f:
            for k, line in enumerate(f, 1):
                if summed.search(line):
                    offenders.append("%s:%d" % (os.path.basename(path), k))
    assert not offenders, offenders
    with open(os.path.join(root, "kernels.hpp")) as f:
        assert "lo + ((hi - lo) >> 1)" in f.read()
    # the midpoint stays inside [lo, hi] for row pointers up to the 32-bit limit
    for lo, hi in ((2 ** 30, 2 ** 30 + 40), (2 ** 31 - 50, 2 ** 31 - 2), (0, 7)):
        mid = lo + ((hi - lo) >> 1)
        assert lo <= mid <= hi and mid < 2 ** 31
        assert lo + hi >= 2 ** 31 or hi < 2 ** 30   # the summed form would have wrapped
