"""Paired solves: two real systems packed as the real and imaginary parts of one complex
vector, solved as two independent real COCG recursions over shared operator and
preconditioner passes (the eigen solver's block-2 inner solves).  Each part must match
the oracle's direct solve of its own system; a zero part must stay exactly zero."""
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.set_device(0)
    from sucemfem.runtime import get_context
    return get_context(0)


def _cavity(order):
    """The shifted PEC cavity operator S - sigma T of a 0.1 m cube (4^3 boxes, two
    materials), its oracle matrix on the free DOFs, and the free DOFs."""
    from sucemfem.Utilities.Workloads import DeviceProblem
    from sucemfem import _hip
    prob = DeviceProblem(4, 0.1, order, _hip.MAT_S | _hip.MAT_T, (1.0, 4.4), seed=1)
    co, ce = O.unit_cube(4, 4, 4)
    osp = O.Space(co * 0.1, ce, order)
    eps = np.array([1.0, 4.4])[prob.region]
    sigma = 1.1 * (np.pi / 0.1) ** 2
    m = O.assemble_fast(osp, eps_r=eps, abc=False)
    free = np.flatnonzero(~osp.pec_mask())
    A = (m["S"] - sigma * m["M"]).tocsc()[free][:, free]
    return prob, A, free, osp.ndofs, sigma


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("precision", [64, 32])
def test_paired_solve_matches_two_direct_solves(dev, order, precision):
    import torch
    from sucemfem import _hip
    _say("paired solve order %d precision %d" % (order, precision))
    prob, A, free, ndofs, sigma = _cavity(order)
    rng = np.random.default_rng(3)
    b0 = np.zeros(ndofs)
    b1 = np.zeros(ndofs)
    b0[free] = rng.standard_normal(len(free))
    b1[free] = 1e3 * rng.standard_normal(len(free))   # parts of different scale
    ref0 = spla.spsolve(A, b0[free])
    ref1 = spla.spsolve(A, b1[free])
    b = torch.as_tensor(b0 + 1j * b1, device="cuda")
    x = torch.zeros_like(b)
    tol = 1e-10
    st = prob.solve_coef(1.0, -sigma, 0.0, b, x, precision=precision, tol=tol,
                         max_iter=20000, check_every=1, precond=_hip.PRECOND_AUXILIARY,
                         eliminate_pec=True, paired=True)
    assert st["status"] == 0 and st["rel_residual"] <= tol, st
    xs = x.cpu().numpy()
    for part, ref, bp in ((xs.real, ref0, b0), (xs.imag, ref1, b1)):
        # each part's own relative residual (the solve reports the larger of the two)
        r = np.linalg.norm(bp[free] - A @ part[free]) / np.linalg.norm(bp[free])
        assert r <= 2 * tol, r
        assert np.linalg.norm(part[free] - ref) / np.linalg.norm(ref) < 1e-7
    pec = np.setdiff1d(np.arange(ndofs), free)
    assert np.abs(xs[pec]).max() == 0.0
    # the same two systems one at a time: the paired recursion runs each real recursion,
    # so it stops within the two single solves' iteration counts (FP64 solves confirm the
    # target with the true residual and restart when it is a rounding above the recursion's,
    # which moves single counts by tens of iterations: 5 % slack on either side)
    its = []
    for bp in (b0, b1):
        xb = torch.zeros_like(b)
        s1 = prob.solve_coef(1.0, -sigma, 0.0, torch.as_tensor(bp + 0j, device="cuda"), xb,
                             precision=precision, tol=tol, max_iter=20000, check_every=1,
                             precond=_hip.PRECOND_AUXILIARY, eliminate_pec=True)
        assert s1["status"] == 0, s1
        its.append(s1["iterations"])
    if precision == 64:
        slack = max(4, max(its) // 20)
        assert min(its) - slack <= st["iterations"] <= max(its) + slack, (st, its)


def test_paired_solve_with_a_zero_part(dev):
    import torch
    from sucemfem import _hip
    _say("paired solve, zero imaginary part")
    prob, A, free, ndofs, sigma = _cavity(2)
    rng = np.random.default_rng(4)
    b0 = np.zeros(ndofs)
    b0[free] = rng.standard_normal(len(free))
    b = torch.as_tensor(b0 + 0j, device="cuda")
    x = torch.zeros_like(b)
    st = prob.solve_coef(1.0, -sigma, 0.0, b, x, precision=64, tol=1e-10, max_iter=20000,
                         check_every=1, precond=_hip.PRECOND_AUXILIARY, eliminate_pec=True,
                         paired=True)
    assert st["status"] == 0, st
    xs = x.cpu().numpy()
    assert np.abs(xs.imag).max() == 0.0
    ref = spla.spsolve(A, b0[free])
    assert np.linalg.norm(xs.real[free] - ref) / np.linalg.norm(ref) < 1e-7


def test_paired_solve_rejects_complex_coefficients(dev):
    import torch
    from sucemfem import _hip
    _say("paired solve rejections")
    prob, A, free, ndofs, sigma = _cavity(1)
    b = torch.zeros(ndofs, dtype=torch.complex128, device="cuda")
    b[0] = 1.0
    x = torch.zeros_like(b)
    with pytest.raises(NotImplementedError):
        prob.solve_coef(1.0, -sigma, 2.0, b, x, precision=64, eliminate_pec=True, paired=True)
    with pytest.raises(NotImplementedError):
        prob.solve_coef(1.0, -sigma, 0.0, b, x, precision=64, eliminate_pec=True, paired=True,
                        method=_hip.METHOD_BICGSTAB, precond=_hip.PRECOND_JACOBI)
