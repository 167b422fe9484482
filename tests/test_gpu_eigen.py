"""GPU cavity eigen solver against the oracle's discrete pencil and the analytic modes."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

DIMS = [1.0, 0.5, 0.75]


def _cavity(order, n=4, dims=DIMS):
    import torch
    torch.cuda.set_device(0)
    from sucemfem.Mesh import Mesh
    from sucemfem.BoundaryConditions import PECWallsBoundaryCondition
    from sucemfem.ProblemConfigurations.EMVectorWaveEigenproblem import EigenProblem
    co, ce = O.box_mesh(dims, n)
    mesh = Mesh(co, ce)
    pec = PECWallsBoundaryCondition()
    pec.init_with_mesh(mesh)
    ep = EigenProblem()
    ep.set_mesh(mesh)
    ep.set_basis_order(order)
    ep.set_boundary_conditions(pec)
    ep.init_problem()
    return ep, (co, ce)


@pytest.mark.parametrize("order", [1, 2])
def test_cavity_k2_matches_oracle_pencil(order):
    """4x4x4 boxes of a 1.0 x 0.5 x 0.75 PEC cavity, sigma = 1.1, 10 lowest k^2."""
    from sucemfem.ProblemConfigurations.EMVectorWaveEigenproblem import DefaultEigenSolver
    print("cavity order %d" % order, flush=True)
    ep, (co, ce) = _cavity(order)
    es = DefaultEigenSolver()
    es.set_eigenproblem(ep)
    es.set_sigma(1.1)
    k2, modes, resid = es.solve_problem_k2(10)
    assert es.stats["status"] == 0, es.stats
    ref, _, _ = O.cavity_eigs(O.Space(co, ce, order), 10, 1.1)
    assert np.all(np.abs(k2 - ref) / ref <= 1e-8), (k2, ref)
    assert np.all(resid < 1e-7)
    ana = O.analytic_cavity_k2(DIMS, 10)
    assert np.all(np.abs(k2 - ana) / ana < (0.08 if order == 1 else 0.007))
    # modes vanish on PEC DOFs
    pec = ep.function_space.pec_mask()
    assert np.abs(modes[:, pec]).max() == 0.0


def test_reference_return_convention():
    """solve_problem returns nu = 1/(k^2 - sigma); the examples recover k^2 = 1/nu + sigma
    (examples/resonant_cavity/driver.py:66)."""
    from sucemfem.ProblemConfigurations.EMVectorWaveEigenproblem import DefaultEigenSolver
    ep, (co, ce) = _cavity(1)
    es = DefaultEigenSolver()
    es.set_eigenproblem(ep)
    es.set_sigma(1.1)
    eigs_w, eigs_v = es.solve_problem(6)
    assert eigs_v.shape == (6, ep.get_global_dimension())
    k2 = np.sort((1 / eigs_w + 1.1).real)
    ref, _, _ = O.cavity_eigs(O.Space(co, ce, 1), 6, 1.1)
    assert np.allclose(k2, ref, rtol=1e-8)


def test_fp32_inner_solves_are_rejected():
    """FP32 inner solves of S - sigma T stall (measured); the solver refuses them."""
    from sucemfem.ProblemConfigurations.EMVectorWaveEigenproblem import DefaultEigenSolver
    ep, _ = _cavity(1)
    es = DefaultEigenSolver()
    es.set_eigenproblem(ep)
    es.set_sigma(1.1)
    es.set_options(inner_precision=32)
    with pytest.raises(NotImplementedError, match="inner_precision 64"):
        es.solve_problem_k2(4)
