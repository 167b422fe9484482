"""Driven problems with essential (PEC) conditions: the reference applies DirichletBC to S,
M and S_0 (SystemMatrices.py:46-58, essential.py:65-81), so the PEC rows of
A = S - k0^2 M + j k0 S_0 are (1 - k0^2 + j k0) e_i while the other rows keep their
couplings to the PEC columns; the RHS is left as assembled (EMDrivenProblem.py:74-80)."""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_dirichlet(A, mask, d):
    A = A.tolil(copy=True)
    for i in np.flatnonzero(mask):
        A.rows[i] = [i]
        A.data[i] = [d]
    return A.tocsr()


@pytest.mark.parametrize("where", ["away", "touching"])
def test_driven_abc_with_interior_pec_sheet(where):
    """ABC on the exterior plus a PEC sheet on the interior plane x = 0.5: the exported LHS
    equals the oracle's LHS with DirichletBC rows, and the device solve (COCG on the free
    DOFs, PEC values moved to the right-hand side) equals the oracle's direct solve."""
    import torch
    torch.cuda.set_device(0)
    from sucemfem.Mesh import Mesh, MeshFunction
    from sucemfem.Sources import current_source, point_source
    from sucemfem.BoundaryConditions import ABCBoundaryCondition, BoundaryConditions
    from sucemfem.BoundaryConditions.essential import EssentialBoundaryCondition
    from sucemfem.ProblemConfigurations.EMDrivenProblem import DrivenProblemABC
    from sucemfem.Utilities.LinalgSolvers import COCGSolver, UMFPACKSolver
    _say("driven ABC + interior PEC sheet, source %s" % where)
    co, ce = O.box_mesh([1.0, 0.5, 0.75], 4)
    osp = O.Space(co, ce, 1)
    faces = osp.topo["faces"]
    sheet = np.all(np.abs(co[faces][:, :, 0] - 0.5) < 1e-12, axis=1)
    assert sheet.sum() == 32
    mesh = Mesh(co, ce)
    mf = MeshFunction("uint", mesh, 2)
    mf.array()[sheet] = 7
    abc = ABCBoundaryCondition()
    abc.set_region_number(1)
    pec = EssentialBoundaryCondition()
    pec.init_with_meshfunction(mf, 7)
    bcs = BoundaryConditions()
    bcs.add_boundary_condition(abc)
    bcs.add_boundary_condition(pec)
    pos = [0.2, 0.2, 0.3] if where == "away" else [0.49, 0.26, 0.4]
    moment = np.array([0.3, -1.0, 2.0])
    srcs = current_source.CurrentSources()
    dip = point_source.PointCurrentSource()
    dip.set_position(pos)
    dip.set_value(moment)
    srcs.add_source(dip)
    freq = 3e8
    dp = DrivenProblemABC()
    dp.set_mesh(mesh)
    dp.set_basis_order(1)
    dp.set_material_regions({0: dict(eps_r=1, mu_r=1)})
    dp.set_region_meshfunction(None)
    dp.set_boundary_conditions(bcs)
    dp.set_sources(srcs)
    dp.set_frequency(freq)
    dp.init_problem()
    A = dp.get_LHS_matrix()
    b = dp.get_RHS()

    k0 = 2 * np.pi * freq / O.c0
    mask = osp.pec_mask(sheet)
    assert np.array_equal(dp.function_space.pec_mask(), mask)
    d = 1 - k0 ** 2 + 1j * k0
    Ao = _oracle_dirichlet(O.driven_lhs(O.assemble(osp), freq), mask, d)
    dn, cv = O.point_source_contrib(osp, pos, moment)
    bo = O.driven_rhs(osp.ndofs, dn, cv, freq)
    assert np.abs(b - bo).max() <= 1e-12 * np.abs(bo).max()
    assert abs(A - Ao).max() <= 1e-12 * abs(Ao).max()
    if where == "touching":
        assert np.abs(bo[mask]).max() > 0       # the source cell has DOFs on the sheet
    else:
        assert np.abs(bo[mask]).max() == 0
    xo = O.solve_direct(Ao, bo)
    solver = COCGSolver(A, 'auxiliary', tol=1e-11, maxiter=20000)
    x = solver.solve(b)
    assert solver.stats["status"] == 0, solver.stats
    rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    assert rel < 1e-8, rel
    assert np.allclose(x[mask], bo[mask] / d, rtol=1e-14, atol=0)
    # the UMFPACK stand-in reaches the same solution
    xu = UMFPACKSolver(A).solve(b)
    assert np.linalg.norm(xu - xo) / np.linalg.norm(xo) < 1e-9
