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


def test_sigma_between_modes_returns_the_modes_above_sigma():
    """sigma between the 3rd and 4th cavity modes: the device returns the 5 lowest k^2
    above sigma, as the oracle's pencil has them (the modes below sigma are not returned:
    the above-sigma contract of sucem_eigs)."""
    from sucemfem.ProblemConfigurations.EMVectorWaveEigenproblem import DefaultEigenSolver
    print("sigma between modes", flush=True)
    ep, (co, ce) = _cavity(2)
    ref_all, _, _ = O.cavity_eigs(O.Space(co, ce, 2), 10, 1.1)
    sigma = 0.5 * (ref_all[2] + ref_all[3])
    es = DefaultEigenSolver()
    es.set_eigenproblem(ep)
    es.set_sigma(sigma)
    k2, _, _ = es.solve_problem_k2(5)
    assert es.stats["status"] == 0, es.stats
    assert np.all(np.abs(k2 - ref_all[3:8]) / ref_all[3:8] <= 1e-8), (k2, ref_all[3:8])


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


@pytest.mark.parametrize("order", [1, 2])
def test_fp32_inner_solves_match_oracle(order):
    """FP32 inner solves (refinement cycles of FP32 auxiliary-space COCG restarted from the
    FP64 residual) give the same 10 lowest k^2 as the oracle's pencil to 1e-8."""
    from sucemfem.ProblemConfigurations.EMVectorWaveEigenproblem import DefaultEigenSolver
    print("cavity order %d, FP32 inner solves" % order, flush=True)
    ep, (co, ce) = _cavity(order)
    es = DefaultEigenSolver()
    es.set_eigenproblem(ep)
    es.set_sigma(1.1)
    es.set_options(inner_precision=32)
    k2, modes, resid = es.solve_problem_k2(10)
    assert es.stats["status"] == 0 and es.stats["inner_failures"] == 0, es.stats
    ref, _, _ = O.cavity_eigs(O.Space(co, ce, order), 10, 1.1)
    assert np.all(np.abs(k2 - ref) / ref <= 1e-8), (k2, ref)
    assert np.all(resid < 1e-7)


def _block_cavity(n, order, seed, inner_precision):
    import torch
    torch.cuda.set_device(0)
    from sucemfem.Utilities.Workloads import dielectric_cavity
    prob = dielectric_cavity(n=n, dims=DIMS, order=order, eps_block=6.0, seed=seed)
    k2, resid, st = prob.eigs(1.1, 20, inner_precision=inner_precision)
    return prob, k2, resid, st


@pytest.mark.parametrize("inner_precision", [64, 32])
def test_dielectric_block_cavity_20_modes(inner_precision):
    """The large eigenproblem's workload at oracle size: the 1.0 x 0.5 x 0.75 PEC cavity with
    a seeded eps_r = 6 block of whole boxes (10-30 % of the volume), order 2, the 20 lowest
    k^2 above sigma = 1.1, against the oracle's shift-invert Lanczos on the same pencil."""
    n = (8, 4, 6)
    print("dielectric block cavity %s, inner FP%d" % (n, inner_precision), flush=True)
    prob, k2, resid, st = _block_cavity(n, 2, 0, inner_precision)
    assert st["status"] == 0 and st["inner_failures"] == 0, st
    co, ce = O.box_mesh(DIMS, n)
    eps = np.where(prob.region == 1, 6.0, 1.0)
    assert 0.1 <= prob.block_fraction <= 0.3 and eps.mean() > 1.0
    osp = O.Space(co, ce, 2)
    assert osp.ndofs == prob.ndofs and osp.ndofs > 4000   # the oracle's eigsh branch
    ref, _, _ = O.cavity_eigs(osp, 20, 1.1, eps_r=eps)
    assert np.all(np.abs(k2 - ref) / ref <= 1e-8), (k2, ref)
    assert np.all(resid < 1e-7), resid
    # the dielectric lowers the spectrum below the empty cavity's
    empty = O.analytic_cavity_k2(DIMS, 1)[0]
    assert k2[0] < empty


def test_homogeneous_eps_scales_device_k2():
    """eps_r = 2.5 in every cell divides every k^2 by 2.5 (closed form: T scales, S not)."""
    import torch
    torch.cuda.set_device(0)
    from sucemfem import _hip
    from sucemfem.Utilities.Workloads import DeviceProblem
    print("homogeneous eps cavity", flush=True)
    which = _hip.MAT_S | _hip.MAT_T
    p1 = DeviceProblem(3, DIMS, 2, which)
    pe = DeviceProblem(3, DIMS, 2, which, eps_values=(2.5, 2.5), seed=1)
    k1, _, s1 = p1.eigs(1.1, 8)
    ke, _, se = pe.eigs(1.1 / 2.5, 8)
    assert s1["status"] == 0 and se["status"] == 0
    assert np.allclose(ke * 2.5, k1, rtol=1e-9, atol=0), (ke * 2.5, k1)


@pytest.mark.parametrize("inner_precision", [64, 32])
def test_relaxed_tolerances_keep_eigenvalues_within_1e8(inner_precision):
    """Lanczos stopped at a Ritz residual of 1e-8 with inner solves to 1e-9 (the serving
    settings): the returned k^2, FP64 Rayleigh quotients of the Ritz vectors, still match
    the oracle's pencil to 1e-8."""
    import torch
    torch.cuda.set_device(0)
    from sucemfem.Utilities.Workloads import dielectric_cavity
    n = (8, 4, 6)
    print("relaxed tolerances, inner FP%d" % inner_precision, flush=True)
    prob = dielectric_cavity(n=n, dims=DIMS, order=2, eps_block=6.0, seed=0)
    k2, resid, st = prob.eigs(1.1, 20, tol=1e-8, inner_tol=1e-9, inner_precision=inner_precision)
    assert st["status"] == 0 and st["inner_failures"] == 0, st
    co, ce = O.box_mesh(DIMS, n)
    eps = np.where(prob.region == 1, 6.0, 1.0)
    ref, _, _ = O.cavity_eigs(O.Space(co, ce, 2), 20, 1.1, eps_r=eps)
    assert np.all(np.abs(k2 - ref) / ref <= 1e-8), (np.abs(k2 - ref) / ref).max()


@pytest.mark.parametrize("order,block,prec", [(1, 1, 64), (2, 1, 64), (1, 2, 64), (2, 2, 64),
                                              (2, 2, 32), (2, 0, 64)])
def test_lanczos_block_sizes_match_oracle(order, block, prec):
    """Block size 1 (one inner solve per basis vector) and 2 (two basis vectors per step,
    their inner solves run as one paired solve) give the oracle pencil's 10 lowest k^2 to
    1e-8; block 0 picks 2 on these meshes (the patch operator serves the paired solves)."""
    from sucemfem.ProblemConfigurations.EMVectorWaveEigenproblem import DefaultEigenSolver
    print("cavity order %d, Lanczos block %d, inner FP%d" % (order, block, prec), flush=True)
    ep, (co, ce) = _cavity(order)
    es = DefaultEigenSolver()
    es.set_eigenproblem(ep)
    es.set_sigma(1.1)
    es.set_options(block=block, inner_precision=prec)
    k2, modes, resid = es.solve_problem_k2(10)
    st = es.stats
    assert st["status"] == 0 and st["inner_failures"] == 0, st
    assert st["block"] == (block or 2), st
    ref, _, _ = O.cavity_eigs(O.Space(co, ce, order), 10, 1.1)
    assert np.all(np.abs(k2 - ref) / ref <= 1e-8), (k2, ref)
    assert np.all(resid < 1e-7)
    # T-orthonormal modes: modes^T T modes = I (T of the oracle, PEC rows excluded)
    m = O.assemble_fast(O.Space(co, ce, order), abc=False)
    gram = modes.real @ (m["M"] @ modes.real.T)
    assert np.abs(gram - np.eye(10)).max() < 1e-7


@pytest.mark.parametrize("nev", [6, 20])
def test_tiny_cavity_basis_capped_by_free_dofs(nev):
    """2x2x2 boxes, order 1: 26 free DOFs (25 nonzero modes), so the default basis is cut
    to the free DOFs (block 2 holds m + 2 orthonormal vectors); the k^2 still match the
    oracle pencil to 1e-8."""
    from sucemfem.ProblemConfigurations.EMVectorWaveEigenproblem import DefaultEigenSolver
    print("tiny cavity, %d modes" % nev, flush=True)
    ep, (co, ce) = _cavity(1, n=2)
    es = DefaultEigenSolver()
    es.set_eigenproblem(ep)
    es.set_sigma(1.1)
    k2, modes, resid = es.solve_problem_k2(nev)
    assert es.stats["status"] == 0, es.stats
    ref, _, _ = O.cavity_eigs(O.Space(co, ce, 1), nev, 1.1)
    assert np.all(np.abs(k2 - ref) / ref <= 1e-8), (k2, ref, es.stats)


@pytest.mark.parametrize("block", [1, 2])
def test_cube_cavity_double_modes(block):
    """The unit cube cavity (4^3 boxes, order 2): its discrete spectrum holds exact double
    eigenvalues (mesh symmetry); block sizes 1 and 2 both return the 14 lowest k^2 with
    their multiplicities, to 1e-8 of the oracle pencil."""
    from sucemfem.ProblemConfigurations.EMVectorWaveEigenproblem import DefaultEigenSolver
    print("cube cavity, block %d" % block, flush=True)
    ep, (co, ce) = _cavity(2, n=4, dims=[1.0, 1.0, 1.0])
    es = DefaultEigenSolver()
    es.set_eigenproblem(ep)
    es.set_sigma(1.1)
    es.set_options(block=block)
    k2, modes, resid = es.solve_problem_k2(14)
    assert es.stats["status"] == 0 and es.stats["block"] == block, es.stats
    ref, _, _ = O.cavity_eigs(O.Space(co, ce, 2), 14, 1.1)
    assert np.sum(np.abs(np.diff(ref)) < 1e-9 * ref[1:]) >= 4   # double modes present
    assert np.all(np.abs(k2 - ref) / ref <= 1e-8), (k2, ref)
