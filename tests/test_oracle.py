"""The FP64 oracle pinned against the reference's golden data and against the mathematics."""
import numpy as np
import pytest
import scipy.linalg

import oracle as O


# ---- reference fixtures ---------------------------------------------------------------
def test_driven_lhs_order1_fixture(golden):
    """LHS_matrix.npy (test_EMDrivenProblem.py:64-72): real part S - k0^2 T, imaginary k0 S0."""
    z = golden("driven_order1_inscribed_tet.npz")
    sp_ = O.Space(*O.inscribed_tet(), 1)
    A = O.driven_lhs(O.assemble(sp_), float(z["frequency"])).toarray()
    assert np.allclose(A, z["lhs"], rtol=1e-10, atol=3e-15)


def test_driven_rhs_order1_fixture(golden):
    """RHS_vector.pickle (test_EMDrivenProblem.py:74-82)."""
    z = golden("driven_order1_inscribed_tet.npz")
    sp_ = O.Space(*O.inscribed_tet(), 1)
    dn, cv = O.point_source_contrib(sp_, z["source_coord"], z["source_value"])
    b = O.driven_rhs(sp_.ndofs, dn, cv, float(z["frequency"]))
    assert np.allclose(b, z["rhs"], rtol=1e-12, atol=1e-16)


def test_filament_order2_fixture(golden):
    """fillament_source_test_data.pickle (test_fillament_source.py:62-69).

    The filament lies on faces shared by two cells; the 26 significant DOFs are
    independent of which cell is used (tangential continuity), the remaining
    DOF ids of the fixture carry |value| < 1e-15.
    """
    f = golden("filament_order2.npz")
    co, ce = O.centred_cube(f["domain_size"], float(f["max_edge_len"]))
    sp_ = O.Space(co, ce, 2)
    dn, v = O.filament_contrib(sp_, f["source_endpoints"], float(f["I"]),
                               int(f["no_integration_points"]))
    mine = dict(zip(dn, v))
    ref = dict(zip(f["dofnos"], f["rhs_contribs"]))
    sig = {k for k, val in ref.items() if abs(val) > 1e-10}
    assert len(sig) == 26
    for k in sig:
        assert abs(mine[k] - ref[k]) <= 1e-12 + 1e-10 * abs(ref[k])
    for k in set(mine) | set(ref):
        if k not in sig:
            assert abs(mine.get(k, 0.0)) < 1e-14 and abs(ref.get(k, 0.0)) < 1e-14


def test_filament_reversal_symmetry(golden):
    """test_fillament_source.py:71-82: reversing the filament negates the contributions."""
    f = golden("filament_order2.npz")
    co, ce = O.centred_cube(f["domain_size"], float(f["max_edge_len"]))
    sp_ = O.Space(co, ce, 2)
    ep = f["source_endpoints"]
    dn, v = O.filament_contrib(sp_, ep, 1.0, 20)
    dn2, v2 = O.filament_contrib(sp_, -ep, 1.0, 20)
    a, b = dict(zip(dn, v)), dict(zip(dn2, v2))
    for k in set(a) | set(b):
        assert abs(a.get(k, 0.) + b.get(k, 0.)) < 1e-13


def _order2_pointsource_system(z, centred=True):
    freq = float(z["freq"])
    lam = O.c0 / freq
    co, ce = O.centred_cube(z["domain_size"], float(z["max_edge_len"]), [0, 0, 0.] if centred else None)
    sp_ = O.Space(co, ce, 2)
    A = O.driven_lhs(O.assemble_fast(sp_), freq)
    dn, cv = O.point_source_contrib(sp_, [0, 0, 0.], np.array([0, 0, 1.]) * lam / 1000)
    return sp_, A, O.driven_rhs(sp_.ndofs, dn, cv, freq)


def test_order2_driven_solution_fixture(golden):
    """reference_dofs-2-...pickle: full order-2 UMFPACK solve (ntff_generate_dofs.py:40-95)."""
    z = golden("driven_order2_dipole_x.npz")
    sp_, A, b = _order2_pointsource_system(z)
    assert sp_.ndofs == 5530
    x = O.solve_direct(A, b)
    assert np.linalg.norm(x - z["x"]) / np.linalg.norm(z["x"]) < 1e-12


def test_order2_filament_solution_fixture(golden):
    """power_flux_reference_dofs_...pickle (1314 DOFs) and its DOLFIN mesh XML."""
    z = golden("driven_order2_filament_x.npz")
    freq = float(z["freq"])
    lam = O.c0 / freq
    co, ce = O.centred_cube(np.array([lam] * 3) * 0.25, lam / 6)
    assert np.array_equal(ce, z["mesh_cells"]) and len(co) == int(z["mesh_nv"])
    sp_ = O.Space(co, ce, 2)
    A = O.driven_lhs(O.assemble_fast(sp_), freq)
    dn, cv = O.filament_contrib(sp_, z["source_endpoints"], float(z["I"]), 100)
    x = O.solve_direct(A, O.driven_rhs(sp_.ndofs, dn, cv, freq))
    assert np.linalg.norm(x - z["x"]) / np.linalg.norm(z["x"]) < 1e-12


@pytest.mark.slow
def test_order2_filament9_fixture_residual(golden):
    """examples/filament_source/data/dofs-2-...-100.pickle came from SciPy's iterative
    solver at its default tolerance: in the oracle system its relative residual is < 1e-5."""
    z = golden("driven_order2_filament9_x.npz")
    co, ce = O.centred_cube(z["domain_size"], float(z["max_edge_len"]))
    sp_ = O.Space(co, ce, 2)
    assert sp_.ndofs == 30186
    A = O.driven_lhs(O.assemble_fast(sp_), float(z["freq"]))
    l = float(z["l"])
    dn, cv = O.filament_contrib(sp_, np.array([[0, 0, -l / 2], [0, 0, l / 2]]), float(z["I"]),
                                int(z["source_integration_points"]))
    b = O.driven_rhs(sp_.ndofs, dn, cv, float(z["freq"]))
    r = np.linalg.norm(A @ z["x"] - b) / np.linalg.norm(b)
    assert r < 1e-5


def test_unitcube1_boundary_edges(golden):
    """test_Geometry.py:31-35: UnitCube(1,1,1) has 18 boundary edges."""
    co, ce = O.unit_cube(1, 1, 1)
    sp_ = O.Space(co, ce, 1)
    assert sp_.pec_mask().sum() == int(golden("unitcube1_topology.npz")["n_boundary_edges"])


# ---- mathematics ----------------------------------------------------------------------
def test_whitney_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(5):
        X = rng.normal(size=(4, 3))
        T, S = O.element_matrices(1, X)
        Tw, Sw = O.whitney_closed_form(X)
        assert np.allclose(T, Tw, rtol=1e-12, atol=1e-14 * abs(Tw).max())
        assert np.allclose(S, Sw, rtol=1e-12, atol=1e-14 * abs(Sw).max())


@pytest.mark.parametrize("order", [1, 2])
def test_basis_is_nodal(order):
    basis, _ = O.reference_basis(order)
    D = O.apply_duals(order, basis)
    assert np.allclose(D, np.eye(len(basis)), atol=1e-13)


def test_numbering_matches_sequential_scan():
    rng = np.random.default_rng(2)
    co, ce = O.unit_cube(3, 2, 2)
    perm = rng.permutation(len(ce))
    ce = ce[perm]
    topo = O.number_entities(ce)
    e, f = O.number_entities_sequential(ce)
    assert np.array_equal(topo["cell_edges"], e)
    assert np.array_equal(topo["cell_faces"], f)


def test_unit_cube_counts():
    """Closed-form entity counts of the Kuhn UnitCube (n=4: 604 edges, 864 faces)."""
    co, ce = O.unit_cube(4, 4, 4)
    t = O.number_entities(ce)
    assert len(t["edges"]) == 604 and len(t["faces"]) == 864 and len(ce) == 384


@pytest.mark.parametrize("order", [1, 2])
def test_curl_of_gradients_vanishes(order):
    """S G = 0 for the discrete gradient; T symmetric positive definite."""
    co, ce = O.box_mesh([1.0, 0.5, 0.75], 2)
    sp_ = O.Space(co, ce, order)
    m = O.assemble_fast(sp_, abc=False)
    G = O.discrete_gradient(sp_)
    assert abs(m["S"] @ G).max() < 1e-12 * abs(m["S"]).max()
    T = m["M"].toarray()
    assert np.allclose(T, T.T, atol=1e-15) and np.linalg.eigvalsh(T).min() > 0


def test_cavity_order1_null_space_and_modes():
    """cfg 4x4x4 box 1.0 x 0.5 x 0.75: null space = interior P1 vertices (27); lowest mode
    within 2% of the analytic TE101 value (order-1 discretisation error)."""
    co, ce = O.box_mesh([1.0, 0.5, 0.75], 4)
    sp_ = O.Space(co, ce, 1)
    m = O.assemble_fast(sp_, abc=False)
    free = np.flatnonzero(~sp_.pec_mask())
    assert sp_.ndofs == 604 and len(free) == 316
    w = scipy.linalg.eigh(m["S"][free][:, free].toarray(), m["M"][free][:, free].toarray(),
                          eigvals_only=True)
    assert (np.abs(w) < 1e-8).sum() == 27
    k2, _, _ = O.cavity_eigs(sp_, 10, 1.1)
    ana = O.analytic_cavity_k2([1.0, 0.5, 0.75], 10)
    assert abs(k2[0] - ana[0]) / ana[0] < 0.02
    assert np.all(np.abs(k2 - ana) / ana < 0.08)


def test_cavity_sigma_between_modes_selects_modes_above_sigma():
    """With sigma between two cavity modes the shift-invert branch (largest positive
    nu = 1/(k^2 - sigma)) returns the nev lowest k^2 above sigma, the same values dense
    eigh of the pencil gives there; the modes below sigma are not returned."""
    co, ce = O.box_mesh([1.0, 0.5, 0.75], 4)
    sp_ = O.Space(co, ce, 1)
    m = O.assemble_fast(sp_, abc=False)
    free = np.flatnonzero(~sp_.pec_mask())
    w = scipy.linalg.eigh(m["S"][free][:, free].toarray(), m["M"][free][:, free].toarray(),
                          eigvals_only=True)
    phys = np.sort(w[w > 1e-6])
    sigma = 0.5 * (phys[2] + phys[3])                     # between the 3rd and 4th modes
    k2, _, _ = O.cavity_eigs(sp_, 5, sigma, dense_limit=0)   # force the eigsh branch
    assert np.allclose(k2, phys[3:8], rtol=1e-10, atol=0), (k2, phys[3:8])
    assert np.all(k2 > sigma)


def test_cavity_order2_close_to_analytic():
    co, ce = O.box_mesh([1.0, 0.5, 0.75], 4)
    sp_ = O.Space(co, ce, 2)
    k2, _, _ = O.cavity_eigs(sp_, 10, 1.1)
    ana = O.analytic_cavity_k2([1.0, 0.5, 0.75], 10)
    assert np.all(np.abs(k2 - ana) / ana < 0.007)


def test_oracle_rejects_point_outside():
    sp_ = O.Space(*O.inscribed_tet(), 1)
    with pytest.raises(AssertionError):
        O.point_source_contrib(sp_, [5, 5, 5.], [1, 0, 0.])


def test_face_matrix_planted_error_is_caught(golden):
    """A plausible mistake (normal component kept in the ABC term) fails the LHS fixture."""
    z = golden("driven_order1_inscribed_tet.npz")
    sp_ = O.Space(*O.inscribed_tet(), 1)
    m = O.assemble(sp_)
    # full N.N instead of the tangential trace
    k = sp_.nloc
    wrong = np.zeros((len(sp_.cells), k, k))
    for c, lf in zip(sp_.boundary_face_cell, sp_.boundary_face_local):
        X = sp_.coords[sp_.cells[c]]
        p, q, r = (X[i] for i in O.LOCAL_FACES[lf])
        nrm = np.cross(q - p, r - p)
        tp, tw = O.tri_rule(4)
        pts = p + np.outer(tp[:, 0], q - p) + np.outer(tp[:, 1], r - p)
        vals, _ = O.eval_basis(1, X, pts)
        wrong[c] += np.einsum("p,pid,pjd->ij", tw * np.linalg.norm(nrm), vals, vals)
    m["S_0"] = O.fem._scatter(sp_, wrong)
    A = O.driven_lhs(m, float(z["frequency"])).toarray()
    assert not np.allclose(A, z["lhs"], rtol=1e-10, atol=3e-15)


def test_cocg_gradient_matches_direct_and_cuts_iterations():
    """Gradient-space preconditioned COCG reaches the direct solution of a heterogeneous
    order-2 driven cube and needs far fewer iterations than Jacobi-COCG; the
    preconditioner is complex symmetric (required by COCG)."""
    co, ce = O.box_mesh([0.1, 0.1, 0.1], 3)
    sp_ = O.Space(co, ce, 2)
    rng = np.random.default_rng(3)
    eps = np.array([1.0, 4.4])[rng.integers(0, 2, len(ce))]
    A = O.driven_lhs(O.assemble_fast(sp_, eps_r=eps), 1e9)
    b = rng.standard_normal(sp_.ndofs) + 1j * rng.standard_normal(sp_.ndofs)
    xd = O.solve_direct(A, b)
    G = O.discrete_gradient(sp_)
    x, its, hist = O.cocg_gradient(A, b, G, tol=1e-10, maxiter=5000)
    assert hist[-1] <= 1e-10
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-7
    _, its_j, _ = O.cocg_jacobi(A, b, tol=1e-10, maxiter=20000)
    assert its * 3 < its_j, (its, its_j)
    gd = O.gradient_diag(A, G)
    assert np.allclose(gd, (G.T @ A @ G).diagonal())
    Minv = (scipy.sparse.diags(1 / A.diagonal()) + G @ scipy.sparse.diags(1 / gd) @ G.T).toarray()
    assert np.allclose(Minv, Minv.T)


@pytest.mark.parametrize("order", [1, 2])
def test_nodal_interpolation_reproduces_fields(order):
    """Pi applied to vertex values of a field the Nedelec space contains (order 2: any
    linear field; order 1: a + b x x) gives DOFs whose basis expansion equals the field
    at random interior points of every cell (an independent check of Pi's points,
    tangents and barycentric weights through the fixture-pinned basis)."""
    co, ce = O.box_mesh([0.3, 0.2, 0.25], 2)
    rng = np.random.default_rng(7)
    inner = np.all((co > 1e-9) & (co < np.array([0.3, 0.2, 0.25]) - 1e-9), axis=1)
    co = co.copy()
    co[inner] += 0.02 * rng.uniform(-1, 1, size=(inner.sum(), 3))
    sp_ = O.Space(co, ce, order)
    a, bvec = rng.standard_normal(3), rng.standard_normal(3)
    M = rng.standard_normal((3, 3))
    if order == 1:
        field = lambda x: a + np.cross(bvec, x)        # noqa: E731
    else:
        field = lambda x: a + x @ M.T                  # noqa: E731
    nodal = np.array([field(x) for x in co]).ravel()
    dofs = O.nodal_interpolation(sp_) @ nodal
    for c in range(len(ce)):
        X = co[ce[c]]
        w = rng.dirichlet(np.ones(4), size=3)
        pts = w @ X
        vals, _ = O.eval_basis(order, X, pts)
        got = np.einsum("pid,i->pd", vals, dofs[sp_.cell_dofs[c]])
        want = np.array([field(x) for x in pts])
        assert np.allclose(got, want, atol=1e-11), (c, got, want)


def test_cocg_auxiliary_matches_direct_and_cuts_iterations():
    """Auxiliary-space COCG (gradient + vector nodal terms) reaches the direct solution of
    a heterogeneous order-2 driven cube in fewer iterations than the gradient form, and
    its preconditioner is complex symmetric."""
    co, ce = O.box_mesh([0.1, 0.1, 0.1], 4)
    sp_ = O.Space(co, ce, 2)
    rng = np.random.default_rng(3)
    eps = np.array([1.0, 4.4])[rng.integers(0, 2, len(ce))]
    A = O.driven_lhs(O.assemble_fast(sp_, eps_r=eps), 1e9)
    b = rng.standard_normal(sp_.ndofs) + 1j * rng.standard_normal(sp_.ndofs)
    xd = O.solve_direct(A, b)
    G = O.discrete_gradient(sp_)
    Pi = O.nodal_interpolation(sp_)
    x, its, hist = O.cocg_auxiliary(A, b, G, Pi, tol=1e-10, maxiter=5000)
    assert hist[-1] <= 1e-10
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-7
    _, its_g, _ = O.cocg_gradient(A, b, G, tol=1e-10, maxiter=5000)
    assert its < its_g, (its, its_g)
    vd = O.gradient_diag(A, Pi)
    Minv = (Pi @ scipy.sparse.diags(1 / vd) @ Pi.T).toarray()
    assert np.allclose(Minv, Minv.T)


# --- post-processing: power flux and near-to-far-field (PostProcessing/tests) ---------
def _flux_space(golden):
    d = golden("driven_order2_filament_x.npz")
    # test_power_flux.py:42 loads the DOLFIN XML mesh, whose coordinates carry 6 digits
    return O.Space(d["mesh_coords"], d["mesh_cells"], 2), d["x"], 2 * np.pi * float(d["freq"]) / O.c0


def test_surface_flux_fixture(golden):
    """``test_power_flux.py:53-54``: SurfaceFlux.calc_flux() == 26.3586862457 (7 places)."""
    sp_, x, k0 = _flux_space(golden)
    assert round(O.surface_flux(sp_, x, k0) - 26.3586862457, 7) == 0


def test_variational_flux_fixture(golden):
    """``test_power_flux.py:72-73``: VariationalSurfaceFlux == 26.9015957282-7.85747240242e-08j."""
    sp_, x, k0 = _flux_space(golden)
    assert round(abs(O.variational_flux(sp_, x, k0) - (26.9015957282 - 7.85747240242e-08j)), 7) == 0


def test_flux_planted_error_is_caught(golden):
    """Dropping the E_i x curl E_r term (or flipping the normal) moves the flux off the fixture."""
    sp_, x, k0 = _flux_space(golden)
    total = 0.0
    for _, w, nrm, E, cE in O.boundary_face_samples(sp_, x):
        total += w @ (np.cross(E.real, -cE.imag) @ nrm)
    assert abs(total / k0 / O.Z0 - 26.3586862457) > 1e-3


def _ntff_space(golden):
    d = golden("driven_order2_dipole_x.npz")
    co, ce = O.centred_cube(d["domain_size"], float(d["max_edge_len"]))
    return O.Space(co, ce, 2), d["x"], float(d["freq"])


def test_surface_ntff_fixture(golden):
    """``test_ntff.py:84-95``: six far-field points, rtol 1e-12, atol 1e-7."""
    sp_, x, f = _ntff_space(golden)
    ff = golden("ntff_surface_order2.npz")
    E = O.surface_ntff(sp_, x, f, ff["theta_deg"], ff["phi_deg"])
    assert np.allclose(E, ff["E_ff"], rtol=1e-12, atol=1e-7)


def test_variational_ntff_fixture(golden):
    """``test_ntff.py:98-103``: variational NTFF, same tolerances."""
    sp_, x, f = _ntff_space(golden)
    ff = golden("ntff_variational_order2.npz")
    E = O.variational_ntff(sp_, x, f, ff["theta_deg"], ff["phi_deg"])
    assert np.allclose(E, ff["E_ff"], rtol=1e-12, atol=1e-7)


def test_ntff_testing_field_fixture(golden):
    """``test_ntff.py:36-55``: (j/k0) rhat x (ahat x rhat) exp(j k0 rhat.x) at 8 points."""
    d = golden("ntff_interpolant.npz")
    for a, want in zip(d["ahats"], d["vals"]):
        got = O.ntff_testing_field(d["rhat"], a, float(d["k0"]))(d["coords"])
        assert np.allclose(got, want, rtol=1e-10, atol=1e-7)


def test_boundary_edge_cells_unit_cube():
    """Cells touching the boundary through an edge: all but the fully interior boxes'
    cells whose edges avoid the boundary; on UnitCube(1) every cell qualifies, and on
    UnitCube(3) none of the central box's cells is missed."""
    co, ce = O.unit_cube(1, 1, 1)
    assert len(O.boundary_edge_cells(O.Space(co, ce, 1))) == 6
    co, ce = O.unit_cube(3, 3, 3)
    sp_ = O.Space(co, ce, 1)
    marked = set(O.boundary_edge_cells(sp_).tolist())
    on_b = lambda p: np.any((np.abs(p) < 1e-12) | (np.abs(p - 1) < 1e-12))  # noqa: E731
    for c in range(len(ce)):
        X = co[ce[c]]
        touches = any(on_b(X[a]) and on_b(X[b]) and
                      np.any((np.abs(X[a] - X[b]) < 1e-12) & ((np.abs(X[a]) < 1e-12) | (np.abs(X[a] - 1) < 1e-12)))
                      for a, b in O.LOCAL_EDGES)
        assert (c in marked) == touches


# --- meshes read from the reference's files ----------------------------------------
def test_facet_region_numbering_matches_dolfin(golden):
    """rectangular_prism_facet_region.xml marks the PEC walls (value 1) by DOLFIN facet
    index; on this unstructured gmsh mesh the first-encounter facet numbering must put
    exactly the 92 boundary facets under those indices."""
    d = golden("mesh_rectangular_prism.npz")
    sp_ = O.Space(d["coords"], d["cells"], 2)
    fr = d["facet_region"]
    assert len(fr) == len(sp_.topo["faces"])
    assert set(np.flatnonzero(fr == 1)) == set(sp_.boundary_faces.tolist())
    assert np.array_equal(sp_.pec_mask(fr == 1), sp_.pec_mask())
    # an interior facet set gives its own closure: 3 edges x 2 + 2 face DOFs each
    inner = np.zeros(len(fr), bool)
    inner[np.flatnonzero(fr == 0)[0]] = True
    assert sp_.pec_mask(inner).sum() == 8


def test_rectangular_prism_cavity_modes(golden):
    """functional_tests/test_EigenProblems.py:65-134 on its own mesh at order 2 (the
    reference uses order 4): the lowest modes of the 1 x 0.5 x 0.25 cavity within the
    order-2 discretisation error of the analytic values."""
    d = golden("mesh_rectangular_prism.npz")
    sp_ = O.Space(d["coords"], d["cells"], 2)
    k2, _, _ = O.cavity_eigs(sp_, 4, 1.1, facet_flags=d["facet_region"] == 1)
    ana = O.analytic_cavity_k2([1.0, 0.5, 0.25], 4)
    assert np.all(np.abs(k2 - ana) / ana < 0.05), (k2, ana)


def test_cavity_homogeneous_eps_scales_k2():
    """A cavity filled with eps_r = e everywhere has k^2 = k^2(vacuum) / e (T scales by e,
    S does not), and 1/mu_r = m everywhere gives m k^2: the per-cell material weights enter
    T and S, and nothing else."""
    co, ce = O.box_mesh([1.0, 0.5, 0.75], 3)
    sp_ = O.Space(co, ce, 2)
    nc = len(ce)
    w1, _, _ = O.cavity_eigs(sp_, 6, 0.5)
    we, _, _ = O.cavity_eigs(sp_, 6, 0.5 / 2.5, eps_r=np.full(nc, 2.5))
    assert np.allclose(we * 2.5, w1, rtol=1e-11, atol=0)
    wm, _, _ = O.cavity_eigs(sp_, 6, 0.5 * 0.4, inv_mu_r=np.full(nc, 0.4))
    assert np.allclose(wm / 0.4, w1, rtol=1e-11, atol=0)
    # a planted mistake (eps applied to S instead of T) would give k^2 e, not k^2 / e
    assert not np.allclose(w1 * 2.5, w1 / 2.5)


def test_dielectric_block_workload_is_seeded():
    """The large-eigenproblem block: box-aligned, 10-30 % of the volume, reproducible."""
    from sucemfem.Utilities.Workloads import dielectric_block
    r1, (lo, hi), f = dielectric_block((120, 60, 90), 0)
    r2, _, _ = dielectric_block((120, 60, 90), 0)
    assert np.array_equal(r1, r2) and 0.1 <= f <= 0.3
    assert r1.mean() == f and len(r1) == 6 * 120 * 60 * 90
    nb = (hi - lo).prod()
    assert r1.sum() == 6 * nb
