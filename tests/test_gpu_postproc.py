"""GPU parity of the boundary post-processing (power flux, NTFF) against the reference's
fixtures (PostProcessing/tests/test_power_flux.py, test_ntff.py) and the oracle."""
import sys

import numpy as np
import pytest

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


def _space(coords, cells, order):
    from sucemfem.Mesh import Mesh, FunctionSpace
    return FunctionSpace(Mesh(np.asarray(coords, np.float64), np.asarray(cells, np.int64)),
                         "Nedelec 1st kind H(curl)", order)


@pytest.fixture(scope="module")
def flux_case(dev, golden):
    d = golden("driven_order2_filament_x.npz")
    V = _space(d["mesh_coords"], d["mesh_cells"], 2)
    return V, d["x"], 2 * np.pi * float(d["freq"]) / O.c0, O.Space(d["mesh_coords"], d["mesh_cells"], 2)


def test_surface_flux_fixture(flux_case):
    """test_power_flux.py:50-54: 26.3586862457 to 7 places; the oracle to 1e-12."""
    from sucemfem.PostProcessing import power_flux
    V, x, k0, osp = flux_case
    _say("surface flux fixture")
    dut = power_flux.SurfaceFlux(V)
    dut.set_k0(k0)
    dut.set_dofs(x)
    f = dut.calc_flux()
    assert round(f - 26.3586862457, 7) == 0, f
    assert abs(f - O.surface_flux(osp, x, k0)) < 1e-12 * abs(f)


def test_variational_flux_fixture(flux_case):
    """test_power_flux.py:69-73: 26.9015957282-7.85747240242e-08j to 7 places."""
    from sucemfem.PostProcessing import power_flux
    V, x, k0, osp = flux_case
    _say("variational flux fixture")
    dut = power_flux.VariationalSurfaceFlux(V)
    dut.set_k0(k0)
    dut.set_dofs(x)
    f = dut.calc_flux()
    assert round(abs(f - (26.9015957282 - 7.85747240242e-08j)), 7) == 0, f
    assert abs(f - O.variational_flux(osp, x, k0)) < 1e-11 * abs(f)


@pytest.fixture(scope="module")
def ntff_case(dev, golden):
    from sucemfem.Utilities.MeshGenerators import get_centred_cube
    d = golden("driven_order2_dipole_x.npz")
    mesh = get_centred_cube(d["domain_size"], float(d["max_edge_len"]))
    from sucemfem.Mesh import FunctionSpace
    V = FunctionSpace(mesh, "Nedelec 1st kind H(curl)", 2)
    co, ce = O.centred_cube(d["domain_size"], float(d["max_edge_len"]))
    return V, d["x"], float(d["freq"]), O.Space(co, ce, 2)


@pytest.mark.parametrize("kind", ["surface", "variational"])
def test_ntff_fixture(ntff_case, golden, kind):
    """test_ntff.py:84-103: six far-field points, rtol 1e-12, atol 1e-7; one call per point
    as the reference's test does, and all points in one call."""
    from sucemfem.PostProcessing import surface_ntff, variational_ntff
    V, x, freq, osp = ntff_case
    _say("ntff fixture %s" % kind)
    ff = golden("ntff_%s_order2.npz" % kind)
    dut = (surface_ntff if kind == "surface" else variational_ntff).NTFF(V)
    dut.set_frequency(freq)
    dut.set_dofs(x)
    E = np.array([dut.calc_pt(t, p) for t, p in zip(ff["theta_deg"], ff["phi_deg"])])
    assert np.allclose(E, ff["E_ff"], rtol=1e-12, atol=1e-7)
    E2 = dut.calc_pts(ff["theta_deg"], ff["phi_deg"])
    assert np.array_equal(E, E2)
    want = (O.surface_ntff if kind == "surface" else O.variational_ntff)(
        osp, x, freq, ff["theta_deg"], ff["phi_deg"])
    assert np.abs(E - want).max() < 1e-10 * np.abs(want).max()


def test_flux_of_device_solution_matches_oracle(dev):
    """Flux and far field of a device-solved heterogeneous driven cube (order 2) agree
    with the oracle evaluated on the same DOF vector; the device vector is used in place."""
    import torch
    from sucemfem.Utilities.Workloads import DeviceProblem
    from sucemfem.PostProcessing import power_flux, surface_ntff
    from sucemfem import _hip
    _say("flux of a device solution")
    prob = DeviceProblem(4, 0.2, 2, _hip.MAT_S | _hip.MAT_T | _hip.MAT_S0, (1.0, 2.2), seed=3)
    k0 = 2 * np.pi * 1e9 / O.c0
    b = prob.rhs_dipoles(np.array([[0.101, 0.099, 0.1]]), np.array([[0, 0, 1e-3]]), k0)
    x = torch.zeros_like(b)
    st = prob.solve(k0, b, x, precision=64, tol=1e-10, max_iter=20000)
    assert st["status"] == 0
    co, ce = O.unit_cube(4, 4, 4)
    osp = O.Space(co * 0.2, ce, 2)
    xh = x.cpu().numpy()
    dut = power_flux.SurfaceFlux(prob.space)
    dut.set_k0(k0)
    dut.set_dofs(x)
    f = dut.calc_flux()
    assert f > 0 and abs(f - O.surface_flux(osp, xh, k0)) < 1e-12 * abs(f)
    nt = surface_ntff.NTFF(prob.space)
    nt.set_frequency(1e9)
    nt.set_dofs(x)
    th, ph = np.array([10.0, 45, 90, 135]), np.array([0.0, 30, 60, 90])
    E = nt.calc_pts(th, ph)
    want = O.surface_ntff(osp, xh, 1e9, th, ph)
    assert np.abs(E - want).max() < 1e-10 * np.abs(want).max()


@pytest.mark.parametrize("order", [1, 2])
def test_reconstruct_points_matches_oracle(dev, order):
    """Reconstruct.reconstruct_points (PostProcessing.py:38-60): the device's point
    location and basis evaluation against the oracle's (lowest-index containing cell,
    fixture-pinned basis) for random complex DOFs at random points, and exact
    reproduction of a field the space contains (order 2: linear; order 1: a + b x x),
    interpolated through Pi, which the reference's own basis pins independently."""
    from sucemfem.PostProcessing import Reconstruct
    _say("reconstruct points order %d" % order)
    co, ce = O.box_mesh([0.3, 0.2, 0.25], 3)
    rng = np.random.default_rng(11)
    inner = np.all((co > 1e-9) & (co < np.array([0.3, 0.2, 0.25]) - 1e-9), axis=1)
    co = co.copy()
    co[inner] += 0.01 * rng.uniform(-1, 1, size=(inner.sum(), 3))
    V = _space(co, ce, order)
    osp = O.Space(co, ce, order)
    assert np.array_equal(V.cell_dofs(), osp.cell_dofs)
    pts = rng.uniform([0.01, 0.01, 0.01], [0.29, 0.19, 0.24], size=(40, 3))
    x = rng.standard_normal(osp.ndofs) + 1j * rng.standard_normal(osp.ndofs)
    dut = Reconstruct(V)
    dut.set_dof_values(x)
    got = dut.reconstruct_points(pts)
    want = np.zeros((len(pts), 3), np.complex128)
    for k, p in enumerate(pts):
        c = O.locate_cell(osp.coords, osp.cells, p)
        vals, _ = O.eval_basis(order, osp.coords[osp.cells[c]], p[None])
        want[k] = vals[0].T @ x[osp.cell_dofs[c]]
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    a, bvec = rng.standard_normal(3), rng.standard_normal(3)
    M = rng.standard_normal((3, 3))
    field = (lambda q: a + np.cross(bvec, q)) if order == 1 else (lambda q: a + q @ M.T)
    dofs = O.nodal_interpolation(osp) @ np.array([field(v) for v in co]).ravel()
    dut.set_dof_values(dofs)
    got = dut.reconstruct_points(pts)
    assert got.dtype == np.float64
    assert np.allclose(got, np.array([field(p) for p in pts]), atol=1e-11)


def test_em_functional_symmetry_and_oracle(dev):
    """CalcEMFunctional (PostProcessing.py:62-172): the reference's symmetry test
    (test_Postprocessing.py:51-60, UnitCube(2,2,2), order 1, random E, g, k0) to rounding,
    the free-space value against the oracle's em_functional, and, with per-cell eps_r and
    mu_r restricted to one cell-domain mark, against the oracle's element matrices summed
    over the marked cells (order 2)."""
    from sucemfem.Mesh import UnitCube, FunctionSpace
    from sucemfem.PostProcessing import CalcEMFunctional
    _say("EM functional")
    rng = np.random.default_rng(5)
    V = FunctionSpace(UnitCube(2, 2, 2), "Nedelec 1st kind H(curl)", 1)
    n = V.dim()
    E = rng.random(n) + 1j * rng.random(n)
    g = rng.random(n) + 1j * rng.random(n)
    k0 = rng.random() * 10
    dut = CalcEMFunctional(V)
    dut.set_k0(k0)
    dut.set_g_dofs(g)
    dut.set_E_dofs(E)
    val1 = dut.calc_functional()
    dut.set_g_dofs(E)
    dut.set_E_dofs(g)
    val2 = dut.calc_functional()
    assert abs(val1 - val2) <= 1e-13 * abs(val1), (val1, val2)
    co, ce = O.unit_cube(2, 2, 2)
    osp = O.Space(co, ce, 1)
    want = O.em_functional(osp, E, g, k0, range(len(ce)))
    assert abs(val1 - want) <= 1e-12 * abs(want), (val1, want)
    # order 2, materials per cell, half the cells marked
    V2 = FunctionSpace(UnitCube(2, 2, 2), "Nedelec 1st kind H(curl)", 2)
    osp2 = O.Space(co, ce, 2)
    nc = len(ce)
    eps = rng.choice([1.0, 2.2, 4.4], nc)
    mu = rng.choice([1.0, 1.5], nc)
    marks = (np.arange(nc) % 2).astype(np.int64)
    E2 = rng.standard_normal(osp2.ndofs) + 1j * rng.standard_normal(osp2.ndofs)
    g2 = rng.standard_normal(osp2.ndofs) + 1j * rng.standard_normal(osp2.ndofs)
    dut = CalcEMFunctional(V2)
    dut.set_k0(k0)
    dut.set_epsr_function(eps)
    dut.set_mur_function(mu)
    dut.set_cell_domains(marks, 1)
    dut.set_E_dofs(E2)
    dut.set_g_dofs(g2)
    got = dut.calc_functional()
    want = 0j
    for c in np.flatnonzero(marks == 1):
        T, S = O.element_matrices(2, co[ce[c]], eps_r=eps[c], inv_mu_r=1.0 / mu[c])
        d = osp2.cell_dofs[c]
        want += g2[d] @ ((S - k0 ** 2 * T) @ E2[d])
    assert abs(got - want) <= 1e-12 * abs(want), (got, want)
