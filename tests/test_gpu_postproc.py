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
