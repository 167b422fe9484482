"""GPU solves on meshes read from the reference's mesh files (DOLFIN XML with facet and
cell regions, .femmesh), against the oracle on the same meshes."""
import numpy as np
import pytest

import oracle as O
from test_mesh_io import write_dolfin_meshfunction, write_dolfin_xml, write_femmesh

pytestmark = pytest.mark.gpu


def _eigs(mesh, order, bc, sigma, nev, materials=None, regions=None):
    import torch
    torch.cuda.set_device(0)
    from sucemfem.ProblemConfigurations.EMVectorWaveEigenproblem import (EigenProblem,
                                                                          DefaultEigenSolver)
    ep = EigenProblem()
    ep.set_mesh(mesh)
    ep.set_basis_order(order)
    ep.set_boundary_conditions(bc)
    if materials is not None:
        ep.set_material_regions(materials)
        ep.set_region_meshfunction(regions)
    ep.init_problem()
    es = DefaultEigenSolver()
    es.set_eigenproblem(ep)
    es.set_sigma(sigma)
    k2, modes, resid = es.solve_problem_k2(nev)
    assert es.stats["status"] == 0, es.stats
    return ep, k2, modes


def test_rectangular_prism_facet_region_pec(tmp_path, golden):
    """functional_tests/test_EigenProblems.py:65-88 at order 2: mesh and PEC facet region
    read from DOLFIN XML, PECWallsBoundaryCondition.init_with_meshfunction(mf, 1)."""
    from sucemfem.Mesh import Mesh, MeshFunction
    from sucemfem.BoundaryConditions import PECWallsBoundaryCondition
    print("rectangular prism facet-region PEC", flush=True)
    d = golden("mesh_rectangular_prism.npz")
    write_dolfin_xml(str(tmp_path / "p.xml"), d["coords"], d["cells"])
    write_dolfin_meshfunction(str(tmp_path / "p_facet_region.xml"), 2, d["facet_region"])
    mesh = Mesh(str(tmp_path / "p.xml"))
    mf = MeshFunction("uint", mesh, str(tmp_path / "p_facet_region.xml"))
    assert mf.dim == 2 and len(mf.array()) == mesh.num_facets() == 362
    bc = PECWallsBoundaryCondition()
    bc.init_with_meshfunction(mf, 1)
    ep, k2, modes = _eigs(mesh, 2, bc, 1.1, 6)
    osp = O.Space(d["coords"], d["cells"], 2)
    flags = d["facet_region"] == 1
    assert np.array_equal(ep.function_space.pec_mask(), osp.pec_mask(flags))
    ref, _, _ = O.cavity_eigs(osp, 6, 1.1, facet_flags=flags)
    assert np.all(np.abs(k2 - ref) / ref <= 1e-8), (k2, ref)
    assert np.abs(modes[:, ep.function_space.pec_mask()]).max() == 0.0


def test_interior_pec_sheet(golden):
    """A PEC sheet on interior facets (essential.py:65-81 with a facet region that is not
    the boundary) combined with the walls: DOF mask and modes match the oracle."""
    from sucemfem.Mesh import Mesh, MeshFunction, boundary_facets
    from sucemfem.BoundaryConditions import PECWallsBoundaryCondition, BoundaryConditions
    from sucemfem.BoundaryConditions.essential import EssentialBoundaryCondition
    print("interior PEC sheet", flush=True)
    co, ce = O.box_mesh([1.0, 0.5, 0.75], 4)
    mesh = Mesh(co, ce)
    osp = O.Space(co, ce, 1)
    # the facets of the plane x = 0.5
    faces = osp.topo["faces"]
    sheet = np.all(np.abs(co[faces][:, :, 0] - 0.5) < 1e-12, axis=1)
    assert sheet.sum() == 32
    mf = MeshFunction("uint", mesh, 2)
    mf.array()[sheet] = 7
    assert set(boundary_facets(mesh).tolist()) == set(osp.boundary_faces.tolist())
    walls = PECWallsBoundaryCondition()
    walls.init_with_mesh(mesh)
    inner = EssentialBoundaryCondition()
    inner.init_with_meshfunction(mf, 7)
    bcs = BoundaryConditions()
    bcs.add_boundary_condition(walls)
    bcs.add_boundary_condition(inner)
    ep, k2, _ = _eigs(mesh, 1, bcs, 1.1, 6)
    flags = sheet.copy()
    flags[osp.boundary_faces] = True
    assert np.array_equal(ep.function_space.pec_mask(), osp.pec_mask(flags))
    ref, _, _ = O.cavity_eigs(osp, 6, 1.1, facet_flags=flags)
    assert np.all(np.abs(k2 - ref) / ref <= 1e-8), (k2, ref)
    # two 0.5 x 0.5 x 0.75 cavities: the lowest mode is TE101 of the half box, doubled
    half = O.analytic_cavity_k2([0.5, 0.5, 0.75], 1)[0]
    assert abs(k2[0] - half) / half < 0.1 and abs(k2[1] - k2[0]) / k2[0] < 0.05


def test_albani_bernardi74_dielectric_cavity(tmp_path, golden):
    """examples/albani_bernardi74/driver.py:53-86 at order 2: mesh and physical regions
    from DOLFIN XML, eps_r 16 / 1, PEC walls, sigma 1.5."""
    from sucemfem.Mesh import Mesh, MeshFunction
    from sucemfem.BoundaryConditions import PECWallsBoundaryCondition
    print("albani-bernardi dielectric cavity", flush=True)
    d = golden("mesh_albani_bernardi74.npz")
    write_dolfin_xml(str(tmp_path / "a.xml"), d["coords"], d["cells"])
    write_dolfin_meshfunction(str(tmp_path / "a_physical_region.xml"), 3, d["physical_region"])
    mesh = Mesh(str(tmp_path / "a.xml"))
    regions = MeshFunction("uint", mesh, str(tmp_path / "a_physical_region.xml"))
    bc = PECWallsBoundaryCondition()
    bc.init_with_mesh(mesh)
    ep, k2, _ = _eigs(mesh, 2, bc, 1.5, 4, {1000: dict(eps_r=16), 1001: dict(eps_r=1)}, regions)
    eps = np.where(d["physical_region"] == 1000, 16.0, 1.0)
    ref, _, _ = O.cavity_eigs(O.Space(d["coords"], d["cells"], 2), 4, 1.5, eps_r=eps)
    assert np.all(np.abs(k2 - ref) / ref <= 1e-8), (k2, ref)


def test_femmesh_sphere_driven(tmp_path, golden):
    """A driven dipole inside the r = 1 m sphere of sandbox/solvers/meshes/sphere-r1m-4.femmesh
    with the ABC on its surface, read through Utilities.MeshIO: the device solve matches the
    oracle's direct solve."""
    import torch
    from sucemfem.Utilities import MeshIO
    from sucemfem.Utilities.LinalgSolvers import solve_sparse_system
    from sucemfem.ProblemConfigurations.EMDrivenProblem import DrivenProblemABC
    from sucemfem.BoundaryConditions import ABCBoundaryCondition, BoundaryConditions
    from sucemfem.Mesh import MeshFunction
    from sucemfem.Sources import point_source, current_source
    print("femmesh sphere driven", flush=True)
    d = golden("mesh_sphere_r1m_4.npz")
    write_femmesh(str(tmp_path / "s.femmesh"), d["nodes"], d["tets"], d["property_nos"])
    mesh = MeshIO.femmesh_2_dolfin_mesh(str(tmp_path / "s.femmesh"))
    freq = 1.0e8
    mf = MeshFunction("uint", mesh, 3)
    mf.set_all(0)
    abc = ABCBoundaryCondition()
    abc.set_region_number(1)
    bcs = BoundaryConditions()
    bcs.add_boundary_condition(abc)
    dp = DrivenProblemABC()
    dp.set_mesh(mesh)
    dp.set_basis_order(1)
    dp.set_material_regions({0: dict(eps_r=1, mu_r=1)})
    dp.set_region_meshfunction(mf)
    dp.set_boundary_conditions(bcs)
    src = point_source.PointCurrentSource()
    src.set_position(np.array([0.05, 0.02, -0.03]))
    src.set_value(np.array([0.0, 0.0, 1.0]))
    cs = current_source.CurrentSources()
    cs.add_source(src)
    dp.set_sources(cs)
    dp.init_problem()
    dp.set_frequency(freq)
    A = dp.get_LHS_matrix()
    b = dp.get_RHS()
    x = solve_sparse_system(A, b, preconditioner_type="diagonal", tol=1e-11)
    osp = O.Space(d["nodes"], d["tets"], 1)
    xo = O.solve_direct(O.driven_lhs(O.assemble_fast(osp), freq), b)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-7


def test_gmsh_rectangular_prism_cavity(tmp_path):
    """rectangular_prism.msh read by Utilities.GmshIO (the gmsh conversion of
    sandbox/pec_labels/pec_labels.py:27-443): its facet_region marks exactly the boundary
    facets, PEC on region 1 gives the oracle's order-2 cavity modes on the same mesh, and
    writing the mesh and facet function to DOLFIN XML and reading them back (the XML path)
    gives the same PEC DOFs and eigenvalues."""
    import os
    from sucemfem.Mesh import Mesh, MeshFunction, boundary_facets
    from sucemfem.Utilities.GmshIO import gmsh_to_dolfin_mesh
    from sucemfem.BoundaryConditions import PECWallsBoundaryCondition
    print("gmsh rectangular prism cavity", flush=True)
    path = os.path.join(os.path.dirname(__file__), "golden", "rectangular_prism.msh")
    mesh, cell_mf, facet_mf = gmsh_to_dolfin_mesh(path)
    assert cell_mf is None and facet_mf is not None
    marked = np.flatnonzero(facet_mf.array() == 1)
    assert np.array_equal(marked, np.sort(boundary_facets(mesh)))
    bc = PECWallsBoundaryCondition()
    bc.init_with_meshfunction(facet_mf, 1)
    ep, k2, _ = _eigs(mesh, 2, bc, 1.1, 5)
    co, ce = mesh.coordinates(), mesh.cells()
    osp = O.Space(co, ce, 2)
    flags = np.zeros(len(facet_mf.array()), bool)
    flags[marked] = True
    assert np.array_equal(ep.function_space.pec_mask(), osp.pec_mask(flags))
    ref, _, _ = O.cavity_eigs(osp, 5, 1.1, facet_flags=flags)
    assert np.all(np.abs(k2 - ref) / ref <= 1e-8), (k2, ref)
    # discretisation check against the analytic modes of the 1 x 0.5 x 0.25 box
    ana = O.analytic_cavity_k2([1.0, 0.5, 0.25], 5)
    assert np.all(np.abs(k2 - ana) / ana < 0.01), (k2, ana)
    write_dolfin_xml(str(tmp_path / "p.xml"), co, ce)
    write_dolfin_meshfunction(str(tmp_path / "p_facet_region.xml"), 2, facet_mf.array())
    mesh2 = Mesh(str(tmp_path / "p.xml"))
    mf2 = MeshFunction("uint", mesh2, str(tmp_path / "p_facet_region.xml"))
    bc2 = PECWallsBoundaryCondition()
    bc2.init_with_meshfunction(mf2, 1)
    ep2, k2b, _ = _eigs(mesh2, 2, bc2, 1.1, 5)
    assert np.array_equal(ep2.function_space.pec_mask(), ep.function_space.pec_mask())
    assert np.all(np.abs(k2b - k2) / k2 <= 1e-12), (k2b, k2)
