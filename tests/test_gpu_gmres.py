"""Device GMRES (LinalgSolvers.GMRESSolver, LinalgSolvers.py:251-260) against the oracle's
direct solve (the reference's UMFPACK path), on the driven operator and on a generic CSR."""
import numpy as np
import pytest
import scipy.sparse

import oracle as O

pytestmark = pytest.mark.gpu


def _driven(order, n=3, freq=3e8):
    import torch
    torch.cuda.set_device(0)
    from sucemfem.Mesh import Mesh
    from sucemfem.Sources import current_source, point_source
    from sucemfem.BoundaryConditions import ABCBoundaryCondition, BoundaryConditions
    from sucemfem.ProblemConfigurations.EMDrivenProblem import DrivenProblemABC
    co, ce = O.box_mesh([1.0, 1.0, 1.0], n)
    abc = ABCBoundaryCondition()
    abc.set_region_number(1)
    bcs = BoundaryConditions()
    bcs.add_boundary_condition(abc)
    srcs = current_source.CurrentSources()
    dip = point_source.PointCurrentSource()
    dip.set_position([0.41, 0.52, 0.47])
    dip.set_value([0.2, -0.3, 1.0])
    srcs.add_source(dip)
    dp = DrivenProblemABC()
    dp.set_mesh(Mesh(co, ce))
    dp.set_basis_order(order)
    dp.set_material_regions({0: dict(eps_r=1, mu_r=1)})
    dp.set_region_meshfunction(None)
    dp.set_boundary_conditions(bcs)
    dp.set_sources(srcs)
    dp.set_frequency(freq)
    dp.init_problem()
    A, b = dp.get_LHS_matrix(), dp.get_RHS()
    osp = O.Space(co, ce, order)
    Ao = O.driven_lhs(O.assemble(osp), freq)
    return A, b, O.solve_direct(Ao, b)


@pytest.mark.parametrize("order,precond,restart", [(1, "diagonal", 20), (2, "auxiliary", 30),
                                                   (2, None, 20)])
def test_gmres_driven_matches_direct(order, precond, restart):
    from sucemfem.Utilities.LinalgSolvers import GMRESSolver
    print("gmres order %d %s restart %d" % (order, precond, restart), flush=True)
    A, b, xo = _driven(order)
    solver = GMRESSolver(A, precond, restart=restart, tol=1e-10, maxiter=20000)
    x = solver.solve(b)
    assert solver.stats["status"] == 0, solver.stats
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1.01e-10
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-7
    # the logged residuals (one per Arnoldi step) reach the tolerance
    res = [r for r in solver.get_logging_data()["res"] if np.isfinite(r)]
    assert len(res) == solver.stats["iterations"] and res[-1] <= 1e-10 * np.linalg.norm(b) * 1.01


def test_gmres_generic_nonsymmetric_csr():
    """A non-symmetric complex matrix (no COCG): GMRES with and without Jacobi."""
    from sucemfem.Utilities.LinalgSolvers import GMRESSolver
    import torch
    torch.cuda.set_device(0)
    rng = np.random.default_rng(4)
    n = 400
    A = scipy.sparse.random(n, n, density=0.02, random_state=5, dtype=np.float64)
    A = A + 1j * scipy.sparse.random(n, n, density=0.01, random_state=6)
    A = (A + scipy.sparse.diags(4.0 + rng.random(n))).tocsr()
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    xo = scipy.sparse.linalg.spsolve(A.tocsc(), b)
    for pre in (None, "diagonal"):
        s = GMRESSolver(A, pre, tol=1e-12, maxiter=5000, restart=15)
        x = s.solve(b)
        assert s.stats["status"] == 0, s.stats
        assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-10
