"""Batched driven sweep (sucem_solve_driven_sweep): many frequencies and dipoles on one mesh
solved together against the oracle's direct solves (the reference's UMFPACK path,
LinalgSolvers.py:262-272, per frequency as in sandbox/patch/patch_driver.py:106-129)."""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _brief(st):
    return [(s["status"], "%.2e" % s["rel_residual"], s["iterations"], s["reliable_updates"],
             s["stagnated"], s["fp64_finish"]) for s in st]


def _oracle_solutions(prob, freqs, pos, mom):
    co, ce = O.box_mesh(prob.dims, prob.n)
    osp = O.Space(co, ce, prob.order)
    eps = np.ones(len(ce)) if prob.region is None else None
    mats = O.assemble(osp)
    out = []
    for f, fr in enumerate(freqs):
        A = O.driven_lhs(mats, fr)
        for s in range(pos.shape[1]):
            dn, cv = O.point_source_contrib(osp, pos[f, s], mom[f, s])
            b = O.driven_rhs(osp.ndofs, dn, cv, fr)
            out.append(O.solve_direct(A, b))
    return np.array(out)


@pytest.mark.parametrize("precision", [64, 32])
def test_sweep_order1_matches_direct(precision):
    """Order 1, 4 frequencies in 0.3-1.5 GHz x 8 seeded unit dipoles (32 systems, batches of
    8 sharing each operator pass, iteration replayed from a hipGraph) against the oracle's
    direct solve of every system."""
    import torch
    torch.cuda.set_device(0)
    from sucemfem import _hip
    from sucemfem.Utilities.Workloads import DeviceProblem, sweep_rhs, sweep_sources
    _say("sweep order 1, FP%d" % precision)
    prob = DeviceProblem(5, 1.0, 1, _hip.MAT_S | _hip.MAT_T | _hip.MAT_S0)
    freqs, pos, mom = sweep_sources(nfreq=4, ndip=8, seed=3)
    B, k0s = sweep_rhs(prob, freqs, pos, mom)
    X = torch.zeros_like(B)
    st = prob.solve_sweep(k0s, B, X, precision=precision, tol=1e-10, check_every=8)
    assert all(s["status"] == 0 and s["rel_residual"] <= 1e-10 for s in st), _brief(st)
    ref = _oracle_solutions(prob, freqs, pos, mom)
    x = X.reshape(len(k0s), -1).cpu().numpy()
    err = np.linalg.norm(x - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert err.max() < 1e-8, err


def test_sweep_graph_width_and_sequential_agree():
    """The same systems with and without the hipGraph (identical launches: bit for bit), in
    batches of 8, 4 and 1 and by the single-system solver (same tolerance)."""
    import torch
    torch.cuda.set_device(0)
    from sucemfem import _hip
    from sucemfem.Utilities.Workloads import DeviceProblem, sweep_rhs, sweep_sources
    _say("sweep graph / width / sequential")
    prob = DeviceProblem(5, 1.0, 1, _hip.MAT_S | _hip.MAT_T | _hip.MAT_S0)
    freqs, pos, mom = sweep_sources(nfreq=2, ndip=4, seed=5)
    B, k0s = sweep_rhs(prob, freqs, pos, mom)
    nb, n = len(k0s), prob.ndofs
    runs = {}
    for key, width, graph in (("g8", 8, True), ("e8", 8, False), ("g4", 4, True),
                              ("g1", 1, True), ("a0", 0, True)):
        X = torch.zeros_like(B)
        st = prob.solve_sweep(k0s, B, X, precision=64, tol=1e-11, width=width, use_graph=graph)
        assert all(s["status"] == 0 for s in st), (key, _brief(st))
        if width:
            assert all(s["batch_width"] == min(width, nb) for s in st), (key, _brief(st))
        else:
            # autotuned: timed groups of 2 and 4 (8 do not fit in the 8 systems), then the
            # faster width for the remaining 2
            assert [s["batch_width"] for s in st[:6]] == [2, 2, 4, 4, 4, 4], _brief(st)
            assert st[6]["batch_width"] == st[7]["batch_width"] == 2
        runs[key] = X.reshape(nb, n).cpu().numpy()
    assert np.array_equal(runs["g8"], runs["e8"])
    for key in ("g4", "g1", "a0"):
        d = np.linalg.norm(runs[key] - runs["g8"], axis=1) / np.linalg.norm(runs["g8"], axis=1)
        assert d.max() < 1e-9, (key, d)
    for j in range(nb):
        x = torch.zeros(n, dtype=torch.complex128, device="cuda")
        st = prob.solve(k0s[j], B.view(nb, n)[j].contiguous(), x, precision=64, tol=1e-11)
        assert st["status"] == 0
        d = np.linalg.norm(x.cpu().numpy() - runs["g8"][j]) / np.linalg.norm(runs["g8"][j])
        assert d < 1e-9, (j, d)


def test_sweep_order2_and_converged_systems_freeze():
    """Order 2 with a per-system spread of difficulty: systems that converge early stop
    moving (their solution is the one whose residual was checked), and a zero right-hand
    side returns zero."""
    import torch
    torch.cuda.set_device(0)
    from sucemfem import _hip
    from sucemfem.Utilities.Workloads import DeviceProblem, sweep_rhs, sweep_sources
    _say("sweep order 2")
    prob = DeviceProblem(3, 1.0, 2, _hip.MAT_S | _hip.MAT_T | _hip.MAT_S0)
    freqs, pos, mom = sweep_sources(nfreq=3, ndip=1, seed=7)
    B, k0s = sweep_rhs(prob, freqs, pos, mom)
    B = torch.cat([B, torch.zeros_like(B[:prob.ndofs])]).contiguous()   # a zero right-hand side
    k0s = np.append(k0s, k0s[0])
    X = torch.zeros_like(B)
    st = prob.solve_sweep(k0s, B, X, precision=32, tol=1e-9, width=4)
    assert all(s["status"] == 0 for s in st), _brief(st)
    x = X.reshape(len(k0s), -1).cpu().numpy()
    assert np.abs(x[-1]).max() == 0.0
    ref = _oracle_solutions(prob, freqs, pos, mom)
    err = np.linalg.norm(x[:-1] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert err.max() < 1e-7, err
    # each system's reported residual is its returned solution's (FP64, recomputed here)
    co, ce = O.box_mesh(prob.dims, prob.n)
    mats = O.assemble(O.Space(co, ce, 2))
    b = B.reshape(len(k0s), -1).cpu().numpy()
    for j in range(3):
        A = O.driven_lhs(mats, freqs[j])
        r = np.linalg.norm(b[j] - A @ x[j]) / np.linalg.norm(b[j])
        assert abs(r - st[j]["rel_residual"]) <= 1e-3 * st[j]["rel_residual"] + 1e-14, (j, r, st[j])
