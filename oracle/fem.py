"""FP64 CPU restatement of the reference's Nedelec hot path (test infrastructure only).

Every function names the reference file:line whose behaviour it restates and
how it is pinned:

=====================================  ===========================================================
function                               pinned by
=====================================  ===========================================================
``unit_cube`` / ``centred_cube``       DOLFIN-written cells of ``power_flux_..._mesh.xml`` (exact);
                                       order-2 driven solution fixture (centring tie-break)
``number_entities`` / ``dofmap``       ``LHS_matrix.npy`` (order 1), ``fillament_source_test_data``
                                       and the three order-2 solution fixtures (order 2); brute-force
                                       sequential first-encounter scan on small meshes
``reference_basis``                    nodal property (dual applied to basis = identity); Whitney
                                       closed form; RHS/filament fixtures
``element_matrices``                   closed-form Whitney mass/stiffness (order 1); ``LHS_matrix.npy``;
                                       invariants S.G = 0 (discrete gradient), T SPD; order-2
                                       solution fixtures
``face_matrices`` (ABC)                imaginary part of ``LHS_matrix.npy``; order-2 solution fixtures
``assemble``                           all of the above; pattern = union of cell pairs
``pec_mask``                           count of boundary entities (closed form on UnitCube)
``point_source_contrib``               ``RHS_vector.pickle`` (exact), filament fixture
``driven_lhs`` / ``driven_rhs``        ``LHS_matrix.npy``, ``RHS_vector.pickle``
``solve_direct``                       order-2 solution fixtures (UMFPACK solves in the reference)
``cavity_eigs``                        analytic cavity modes (discretisation check), dense ``eigh``
                                       on the eliminated pencil (small), null-space dimension
``discrete_gradient``                  S.G = 0; G applied to nodal values of a linear function
                                       reproduces its gradient's DOFs
``nodal_interpolation``                reproduction of linear (order 2) / a + b x x (order 1)
                                       fields through the basis expansion
``cocg_gradient`` / ``cocg_auxiliary`` convergence to ``solve_direct``'s solution
=====================================  ===========================================================
"""
from __future__ import division

import itertools
import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# ---------------------------------------------------------------------------
# Constants -- sucemfem/Consts.py:24-33 (evaluated in the same order)
# ---------------------------------------------------------------------------
mu0 = np.pi * 4 * 1e-7
c0 = 299792458
eps0 = 10 ** 7 / 4 / np.pi / c0 / c0
Z0 = mu0 * c0

# UFC local entities of a tetrahedron with sorted vertices
LOCAL_EDGES = np.array([(2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1)], dtype=np.int64)
LOCAL_FACES = np.array([(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)], dtype=np.int64)
REF_VERTS = np.array([[0., 0., 0.], [1., 0., 0.], [0., 1., 0.], [0., 0., 1.]])


# ---------------------------------------------------------------------------
# Meshes
# ---------------------------------------------------------------------------
def unit_cube(nx, ny, nz):
    """DOLFIN 1.0 ``UnitCube(nx, ny, nz)``: Kuhn split of each box into 6 tets.

    Follows the DOLFIN mesh that ``Utilities/MeshGenerators.py:33`` builds.
    Vertex ``v = iz*(nx+1)*(ny+1) + iy*(nx+1) + ix`` at ``(ix/nx, iy/ny, iz/nz)``;
    boxes visited with ix fastest; each cell's vertex list sorted ascending
    (UFC ordering).  Returns ``(coords float64[nv,3], cells int64[nc,4])``.
    """
    ix, iy, iz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    ix, iy, iz = (a.transpose(2, 1, 0).ravel() for a in (ix, iy, iz))
    coords = np.stack([ix / nx, iy / ny, iz / nz], axis=1).astype(np.float64)
    bx, by, bz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    bx, by, bz = (a.transpose(2, 1, 0).ravel() for a in (bx, by, bz))
    sx, sxy = nx + 1, (nx + 1) * (ny + 1)
    v0 = bz * sxy + by * sx + bx
    v1 = v0 + 1
    v2 = v0 + sx
    v3 = v1 + sx
    v4, v5, v6, v7 = v0 + sxy, v1 + sxy, v2 + sxy, v3 + sxy
    six = np.stack([
        np.stack([v0, v1, v3, v7], 1), np.stack([v0, v1, v7, v5], 1),
        np.stack([v0, v5, v7, v4], 1), np.stack([v0, v3, v2, v7], 1),
        np.stack([v0, v6, v4, v7], 1), np.stack([v0, v2, v6, v7], 1)], axis=1)
    cells = np.sort(six.reshape(-1, 4), axis=1).astype(np.int64)
    return coords, cells


def box_mesh(dims, n):
    """UnitCube(n) scaled per axis by ``dims`` (as ``examples/resonant_cavity/driver.py:34-40``)."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    coords, cells = unit_cube(nx, ny, nz)
    coords = coords * np.asarray(dims, dtype=np.float64)[None, :]
    return coords, cells


def cell_midpoints(coords, cells):
    return coords[cells].mean(axis=1)


def centred_cube(domain_size, max_edge_len, centred_element_coordinate=None):
    """``get_centred_cube`` -- ``sucemfem/Utilities/MeshGenerators.py:26-53``.

    n = ceil(sqrt(2)*L/h) boxes per axis; coordinates scaled by L and shifted by
    -L/2.  With ``centred_element_coordinate`` the mesh is translated so that the
    point is at the centroid of the containing cell whose midpoint is nearest
    (ties -> lowest cell index, i.e. the first Kuhn tet of a box).
    """
    domain_size = np.asarray(domain_size, dtype=np.float64) * np.ones(3)
    n = np.array(np.ceil(np.sqrt(2) * domain_size / max_edge_len), dtype=np.int64)
    coords, cells = unit_cube(*n)
    coords = coords * domain_size
    coords = coords - domain_size / 2
    if centred_element_coordinate is not None:
        p = np.asarray(centred_element_coordinate, dtype=np.float64)
        cand = containing_cells(coords, cells, p)
        mids = cell_midpoints(coords, cells[cand])
        d = np.linalg.norm(mids - p, axis=1)
        best = cand[np.flatnonzero(d <= d.min() * (1 + 1e-12) + 1e-300)[0]]
        coords = coords - cell_midpoints(coords, cells[best:best + 1])[0]
    return coords, cells


def inscribed_tet():
    """``InscribedTet`` -- ``sucemfem/Testing/Meshes.py:25-61`` (cells sorted as DOLFIN orders them)."""
    coords = np.array([[-1. / 2, 1. / 2, -1. / 2], [1. / 2, 1. / 2, 1. / 2],
                       [1. / 2, -1. / 2, -1. / 2], [-1. / 2, -1. / 2, 1. / 2],
                       [1. / 6, 1. / 6, -1. / 6], [-1. / 6, 1. / 6, 1. / 6],
                       [-1. / 6, -1. / 6, -1. / 6], [1. / 6, -1. / 6, 1. / 6]], np.float64)
    cells = np.array([[1, 4, 6, 7], [1, 3, 5, 7], [3, 4, 7, 8], [1, 2, 5, 6], [2, 3, 5, 8],
                      [2, 4, 6, 8], [4, 6, 7, 8], [1, 5, 6, 7], [3, 5, 7, 8], [2, 5, 6, 8],
                      [5, 6, 7, 8]], np.int64) - 1
    return coords, np.sort(cells, axis=1)


# ---------------------------------------------------------------------------
# Topology and DOF numbering (DOLFIN first encounter)
# ---------------------------------------------------------------------------
def _first_encounter(keys):
    """Global ids of the rows of ``keys`` (n, k) numbered in order of first appearance."""
    uniq, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty(len(uniq), dtype=np.int64)
    rank[order] = np.arange(len(uniq))
    return uniq[order], rank[inv.ravel()]


def number_entities(cells):
    """Edges and faces numbered by first encounter (scan cells, then local entities).

    Restates the DOLFIN topology computation behind ``dolfin.FunctionSpace`` at
    ``ProblemConfigurations/EMProblem.py:100-105``.  Returns a dict with
    ``edges`` (nE,2), ``cell_edges`` (nc,6), ``faces`` (nF,3), ``cell_faces`` (nc,4)
    and ``face_cells`` (nF,2; -1 where a face has a single cell).
    """
    cells = np.asarray(cells, dtype=np.int64)
    nc = len(cells)
    ek = cells[:, LOCAL_EDGES].reshape(-1, 2)
    edges, eid = _first_encounter(ek)
    fk = cells[:, LOCAL_FACES].reshape(-1, 3)
    faces, fid = _first_encounter(fk)
    cell_faces = fid.reshape(nc, 4)
    face_cells = -np.ones((len(faces), 2), dtype=np.int64)
    owner = np.repeat(np.arange(nc), 4)
    lf = np.tile(np.arange(4), nc)
    order = np.lexsort((owner, fid))
    fid_s, own_s, lf_s = fid[order], owner[order], lf[order]
    first = np.ones(len(fid_s), bool)
    first[1:] = fid_s[1:] != fid_s[:-1]
    face_cells[fid_s[first], 0] = own_s[first]
    face_cells[fid_s[~first], 1] = own_s[~first]
    face_local = -np.ones((len(faces), 2), dtype=np.int64)
    face_local[fid_s[first], 0] = lf_s[first]
    face_local[fid_s[~first], 1] = lf_s[~first]
    return dict(edges=edges, cell_edges=eid.reshape(nc, 6), faces=faces,
                cell_faces=cell_faces, face_cells=face_cells, face_local=face_local)


def number_entities_sequential(cells):
    """Literal sequential first-encounter scan (used to pin ``number_entities``)."""
    emap, fmap = {}, {}
    ce = np.zeros((len(cells), 6), np.int64)
    cf = np.zeros((len(cells), 4), np.int64)
    for c, cell in enumerate(cells):
        for l, (a, b) in enumerate(LOCAL_EDGES):
            k = (cell[a], cell[b])
            if k not in emap:
                emap[k] = len(emap)
            ce[c, l] = emap[k]
        for l, f in enumerate(LOCAL_FACES):
            k = tuple(cell[f])
            if k not in fmap:
                fmap[k] = len(fmap)
            cf[c, l] = fmap[k]
    return ce, cf


def dofmap(order, topo):
    """Cell -> global DOF map for Nedelec order 1 or 2.

    Order 1: DOF = global edge id.  Order 2: local DOFs are 2 per local edge
    (points 1/3 then 2/3 from the lower vertex) then 2 per local face; global
    edge DOF ``2*E+k``, face DOF ``2*nE + 2*F + k``.
    Returns ``(cell_dofs int64[nc, 6|20], ndofs)``.
    """
    ce, cf = topo["cell_edges"], topo["cell_faces"]
    nE, nF = len(topo["edges"]), len(topo["faces"])
    if order == 1:
        return ce.copy(), nE
    if order == 2:
        nc = len(ce)
        d = np.empty((nc, 20), np.int64)
        d[:, 0:12:2] = 2 * ce
        d[:, 1:12:2] = 2 * ce + 1
        d[:, 12::2] = 2 * nE + 2 * cf
        d[:, 13::2] = 2 * nE + 2 * cf + 1
        return d, 2 * nE + 2 * nF
    raise NotImplementedError("only orders 1 and 2")


# ---------------------------------------------------------------------------
# Polynomial machinery on the reference tetrahedron
# ---------------------------------------------------------------------------
MONO = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (1, 0, 1),
        (0, 2, 0), (0, 1, 1), (0, 0, 2)]
_MIDX = {m: i for i, m in enumerate(MONO)}


def _mono_vals(pts):
    pts = np.atleast_2d(pts)
    return np.stack([pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c for a, b, c in MONO], axis=1)


def _times_coord(coef, axis):
    """Multiply a scalar polynomial (coefficients over MONO) by x_axis."""
    out = np.zeros_like(coef)
    for i, m in enumerate(MONO):
        if coef[i] != 0:
            mm = list(m)
            mm[axis] += 1
            out[_MIDX[tuple(mm)]] += coef[i]
    return out


def _deriv(coef, axis):
    out = np.zeros_like(coef)
    for i, m in enumerate(MONO):
        if coef[i] != 0 and m[axis] > 0:
            mm = list(m)
            mm[axis] -= 1
            out[_MIDX[tuple(mm)]] += coef[i] * m[axis]
    return out


def _cross_x(v):
    """x cross v for a vector polynomial v (3, nmono)."""
    return np.stack([_times_coord(v[2], 1) - _times_coord(v[1], 2),
                     _times_coord(v[0], 2) - _times_coord(v[2], 0),
                     _times_coord(v[1], 0) - _times_coord(v[0], 1)])


def _curl(v):
    return np.stack([_deriv(v[2], 1) - _deriv(v[1], 2),
                     _deriv(v[0], 2) - _deriv(v[2], 0),
                     _deriv(v[1], 0) - _deriv(v[0], 1)])


def _eval_field(v, pts):
    return _mono_vals(pts) @ v.T           # (npts, 3)


def _dual_functionals(order):
    """(point, tangent) pairs of the FIAT nodal dual, in local DOF order."""
    funcs = []
    for a, b in LOCAL_EDGES:
        t = REF_VERTS[b] - REF_VERTS[a]
        svals = [0.5] if order == 1 else [1. / 3, 2. / 3]
        for s in svals:
            funcs.append((REF_VERTS[a] + s * t, t))
    if order == 2:
        for p, q, r in LOCAL_FACES:
            cen = (REF_VERTS[p] + REF_VERTS[q] + REF_VERTS[r]) / 3
            funcs.append((cen, REF_VERTS[q] - REF_VERTS[p]))
            funcs.append((cen, REF_VERTS[r] - REF_VERTS[p]))
    return funcs


_BASIS_CACHE = {}


def reference_basis(order):
    """Nodal Nedelec-1 basis on the reference tet as vector polynomials (ndof, 3, nmono).

    Space: P_{k-1}^3 + x cross (homogeneous P_{k-1})^3; nodal w.r.t. tangential
    point evaluations with unnormalised edge/face tangents (FIAT convention).
    """
    if order in _BASIS_CACHE:
        return _BASIS_CACHE[order]
    nm = len(MONO)
    prim = []
    unit = np.eye(nm)
    if order == 1:
        for a in range(3):
            v = np.zeros((3, nm)); v[a, 0] = 1; prim.append(v)
        for a in range(3):
            v = np.zeros((3, nm)); v[a, 0] = 1; prim.append(_cross_x(v))
    elif order == 2:
        for a in range(3):
            for m in range(4):
                v = np.zeros((3, nm)); v[a, m] = 1; prim.append(v)
        for a in range(3):
            for bb in range(3):
                if (a, bb) == (2, 2):
                    continue
                v = np.zeros((3, nm)); v[a] = unit[_MIDX[tuple(np.eye(3, dtype=int)[bb])]]
                prim.append(_cross_x(v))
    else:
        raise NotImplementedError
    prim = np.array(prim)
    duals = _dual_functionals(order)
    V = np.array([[_eval_field(phi, pt[None])[0] @ t for phi in prim] for pt, t in duals])
    C = np.linalg.solve(V, np.eye(len(prim)))
    basis = np.einsum("kj,kdm->jdm", C, prim)
    curls = np.array([_curl(b) for b in basis])
    _BASIS_CACHE[order] = (basis, curls)
    return basis, curls


def apply_duals(order, fields):
    """Apply the nodal dual to vector polynomials (n, 3, nmono) -> (n, ndof)."""
    duals = _dual_functionals(order)
    return np.array([[_eval_field(f, pt[None])[0] @ t for pt, t in duals] for f in fields])


# ---------------------------------------------------------------------------
# Quadrature (collapsed Gauss-Legendre, exact for the degrees used here)
# ---------------------------------------------------------------------------
def tet_rule(n=4):
    x, w = np.polynomial.legendre.leggauss(n)
    x = (x + 1) / 2; w = w / 2
    pts, wts = [], []
    for i, j, k in itertools.product(range(n), repeat=3):
        a, b, c = x[i], x[j], x[k]
        # Duffy: (a,b,c) in unit cube -> tet
        px = a * (1 - b) * (1 - c)
        py = b * (1 - c)
        pz = c
        jac = (1 - b) * (1 - c) ** 2
        pts.append((px, py, pz)); wts.append(w[i] * w[j] * w[k] * jac)
    return np.array(pts), np.array(wts)


def tri_rule(n=4):
    x, w = np.polynomial.legendre.leggauss(n)
    x = (x + 1) / 2; w = w / 2
    pts, wts = [], []
    for i, j in itertools.product(range(n), repeat=2):
        a, b = x[i], x[j]
        pts.append((a * (1 - b), b)); wts.append(w[i] * w[j] * (1 - b))
    return np.array(pts), np.array(wts)


# ---------------------------------------------------------------------------
# Physical basis evaluation and element matrices
# ---------------------------------------------------------------------------
def cell_geometry(X):
    """Jacobian J (columns X1-X0, X2-X0, X3-X0), its determinant and inverse."""
    J = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], axis=1)
    return J, np.linalg.det(J), np.linalg.inv(J)


def eval_basis(order, X, pts):
    """Physical basis values and curls at physical points of the cell with vertices X.

    Covariant Piola map N = J^-T N^(x^), curl N = J curl N^ / det J.
    Returns ``(vals (npts, ndof, 3), curls (npts, ndof, 3))``.
    """
    basis, curls = reference_basis(order)
    J, detJ, Jinv = cell_geometry(X)
    xh = (np.atleast_2d(pts) - X[0]) @ Jinv.T
    m = _mono_vals(xh)
    vh = np.einsum("pm,idm->pid", m, basis)
    ch = np.einsum("pm,idm->pid", m, curls)
    vals = vh @ Jinv            # (J^-T v)_a = sum_b Jinv[b,a] v_b
    cvals = ch @ J.T / detJ
    return vals, cvals


def element_matrices(order, X, eps_r=1.0, inv_mu_r=1.0):
    """Element mass T_e = int eps_r N_i.N_j and stiffness S_e = int (1/mu_r) curl N_i . curl N_j.

    ``sucemfem/Forms.py:71-85`` (UFL forms integrated exactly by FFC).
    """
    qp, qw = tet_rule(4)
    J, detJ, _ = cell_geometry(X)
    pts = X[0] + qp @ J.T
    vals, cvals = eval_basis(order, X, pts)
    w = qw * abs(detJ)
    T = eps_r * np.einsum("p,pid,pjd->ij", w, vals, vals)
    S = inv_mu_r * np.einsum("p,pid,pjd->ij", w, cvals, cvals)
    return T, S


def whitney_closed_form(X):
    """Textbook closed-form order-1 (Whitney) mass and stiffness, used to pin ``element_matrices``."""
    J, detJ, Jinv = cell_geometry(X)
    V = abs(detJ) / 6
    # barycentric gradients: lambda_0 = 1 - sum, lambda_k = xhat_k
    g = np.vstack([-Jinv.sum(axis=0), Jinv])        # (4,3)
    T = np.zeros((6, 6)); S = np.zeros((6, 6))
    for a, (i, j) in enumerate(LOCAL_EDGES):
        for b, (k, l) in enumerate(LOCAL_EDGES):
            d = lambda p, q: 1.0 if p == q else 0.0
            T[a, b] = V / 20 * (g[j] @ g[l] * (1 + d(i, k)) - g[j] @ g[k] * (1 + d(i, l))
                                - g[i] @ g[l] * (1 + d(j, k)) + g[i] @ g[k] * (1 + d(j, l)))
            S[a, b] = V * 4 * np.cross(g[i], g[j]) @ np.cross(g[k], g[l])
    return T, S


def face_matrix(order, X, local_face):
    """ABC face term int_F (n x N_i).(n x N_j) dS on local face ``local_face`` of the cell.

    ``sucemfem/BoundaryConditions/ABC.py:56-57``; full ndof x ndof (zero off the face DOFs).
    """
    p, q, r = (X[i] for i in LOCAL_FACES[local_face])
    nrm = np.cross(q - p, r - p)
    area2 = np.linalg.norm(nrm)
    n = nrm / area2
    tp, tw = tri_rule(4)
    pts = p + np.outer(tp[:, 0], q - p) + np.outer(tp[:, 1], r - p)
    vals, _ = eval_basis(order, X, pts)
    vt = vals - np.einsum("pid,d->pi", vals, n)[:, :, None] * n[None, None, :]
    return np.einsum("p,pid,pjd->ij", tw * area2, vt, vt)


# ---------------------------------------------------------------------------
# Space, assembly and boundary data
# ---------------------------------------------------------------------------
class Space(object):
    """Oracle function space: mesh + order + numbering + pattern."""

    def __init__(self, coords, cells, order):
        self.coords = np.asarray(coords, dtype=np.float64)
        self.cells = np.sort(np.asarray(cells, dtype=np.int64), axis=1)
        self.order = order
        self.topo = number_entities(self.cells)
        self.cell_dofs, self.ndofs = dofmap(order, self.topo)
        fc = self.topo["face_cells"]
        self.boundary_faces = np.flatnonzero(fc[:, 1] < 0)
        self.boundary_face_cell = fc[self.boundary_faces, 0]
        self.boundary_face_local = self.topo["face_local"][self.boundary_faces, 0]

    @property
    def nloc(self):
        return self.cell_dofs.shape[1]

    def pattern(self):
        """DOLFIN sparsity pattern: union of all cell-local DOF pairs, sorted columns."""
        d = self.cell_dofs
        k = d.shape[1]
        rows = np.repeat(d, k, axis=1).ravel()
        cols = np.tile(d, (1, k)).ravel()
        A = sp.coo_matrix((np.zeros(len(rows)), (rows, cols)), shape=(self.ndofs,) * 2).tocsr()
        A.sum_duplicates()
        A.sort_indices()
        return A.indptr.astype(np.int64), A.indices.astype(np.int64)

    def face_dofs_local(self, local_face):
        """Local DOF indices whose tangential trace lives on local face ``local_face``."""
        fverts = set(LOCAL_FACES[local_face])
        loc = []
        for l, (a, b) in enumerate(LOCAL_EDGES):
            if a in fverts and b in fverts:
                loc += [l] if self.order == 1 else [2 * l, 2 * l + 1]
        if self.order == 2:
            loc += [12 + 2 * local_face, 13 + 2 * local_face]
        return np.array(loc, np.int64)

    def pec_mask(self, facet_flags=None):
        """DOFs on the closure (edges and faces) of the flagged facets; default: the
        boundary facets.

        ``BoundaryConditions/essential.py:86-99`` marks all boundary facets, and
        ``essential.py:65-81`` takes a facet MeshFunction and region number; the
        topological DirichletBC constrains the DOFs of the closure.
        """
        mask = np.zeros(self.ndofs, bool)
        if facet_flags is None:
            pairs = zip(self.boundary_face_cell, self.boundary_face_local)
        else:
            flagged = np.flatnonzero(np.asarray(facet_flags))
            fc, fl = self.topo["face_cells"], self.topo["face_local"]
            pairs = [(fc[f, s], fl[f, s]) for f in flagged for s in (0, 1) if fc[f, s] >= 0]
        for c, lf in pairs:
            mask[self.cell_dofs[c, self.face_dofs_local(lf)]] = True
        return mask


def _scatter(space, rows_local_vals):
    d = space.cell_dofs
    k = d.shape[1]
    rows = np.repeat(d, k, axis=1).ravel()
    cols = np.tile(d, (1, k)).ravel()
    A = sp.coo_matrix((rows_local_vals.ravel(), (rows, cols)), shape=(space.ndofs,) * 2).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return A


def assemble(space, eps_r=None, inv_mu_r=None, abc=True):
    """Assemble ``{'M': T, 'S': S, 'S_0': S0}`` as scipy CSR on the DOLFIN pattern.

    ``SystemMatrices.py:46-58`` + ``EMDrivenProblem.py:33-38``.  Explicit zeros
    of the pattern are kept (as the uBLAS export at ``Converters.py:27-44``).
    """
    nc = len(space.cells)
    eps_r = np.ones(nc) if eps_r is None else np.broadcast_to(np.asarray(eps_r, float), (nc,))
    inv_mu_r = np.ones(nc) if inv_mu_r is None else np.broadcast_to(np.asarray(inv_mu_r, float), (nc,))
    k = space.nloc
    Te = np.empty((nc, k, k)); Se = np.empty((nc, k, k))
    for c in range(nc):
        Te[c], Se[c] = element_matrices(space.order, space.coords[space.cells[c]], eps_r[c], inv_mu_r[c])
    out = {"M": _scatter(space, Te), "S": _scatter(space, Se)}
    if abc:
        S0e = np.zeros((nc, k, k))
        for c, lf in zip(space.boundary_face_cell, space.boundary_face_local):
            S0e[c] += face_matrix(space.order, space.coords[space.cells[c]], lf)
        out["S_0"] = _scatter(space, S0e)
    return out


def element_matrices_batched(order, coords, cells, eps_r, inv_mu_r):
    """Vectorised element matrices (same mathematics as ``element_matrices``), for larger meshes."""
    basis, curls = reference_basis(order)
    qp, qw = tet_rule(4)
    m = _mono_vals(qp)
    vh = np.einsum("pm,idm->pid", m, basis)          # (q, k, 3)
    ch = np.einsum("pm,idm->pid", m, curls)
    X = coords[cells]
    J = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)
    detJ = np.linalg.det(J)
    Jinv = np.linalg.inv(J)
    G = np.einsum("cab,ceb->cae", Jinv, Jinv)         # Jinv Jinv^T
    K = np.einsum("cba,cbe->cae", J, J)               # J^T J
    R = np.einsum("q,qia,qjb->abij", qw, vh, vh)
    Q = np.einsum("q,qia,qjb->abij", qw, ch, ch)
    T = np.einsum("c,cab,abij->cij", eps_r * np.abs(detJ), G, R)
    S = np.einsum("c,cab,abij->cij", inv_mu_r / np.abs(detJ), K, Q)
    return T, S


def assemble_fast(space, eps_r=None, inv_mu_r=None, abc=True):
    """Same result as ``assemble`` with batched element matrices (face terms per face)."""
    nc = len(space.cells)
    eps_r = np.ones(nc) if eps_r is None else np.broadcast_to(np.asarray(eps_r, float), (nc,))
    inv_mu_r = np.ones(nc) if inv_mu_r is None else np.broadcast_to(np.asarray(inv_mu_r, float), (nc,))
    Te, Se = element_matrices_batched(space.order, space.coords, space.cells, eps_r, inv_mu_r)
    out = {"M": _scatter(space, Te), "S": _scatter(space, Se)}
    if abc:
        k = space.nloc
        S0e = np.zeros((nc, k, k))
        for c, lf in zip(space.boundary_face_cell, space.boundary_face_local):
            S0e[c] += face_matrix(space.order, space.coords[space.cells[c]], lf)
        out["S_0"] = _scatter(space, S0e)
    return out


# ---------------------------------------------------------------------------
# Sources and the driven system
# ---------------------------------------------------------------------------
def barycentric(coords, cells, p):
    X = coords[cells]
    J = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)
    lam = np.linalg.solve(J, (p - X[:, 0])[:, :, None])[:, :, 0]
    return np.concatenate([1 - lam.sum(axis=1, keepdims=True), lam], axis=1)


def containing_cells(coords, cells, p, tol=1e-12):
    lam = barycentric(coords, cells, np.asarray(p, float))
    return np.flatnonzero(np.all(lam >= -tol, axis=1))


def locate_cell(coords, cells, p, tol=1e-12):
    """Lowest-index cell containing ``p`` (the reference asserts that one exists,
    ``Sources/point_source.py:79``)."""
    cand = containing_cells(coords, cells, p, tol)
    if len(cand) == 0:
        raise AssertionError("point %r is outside the mesh" % (p,))
    return int(cand[0])


def point_source_contrib(space, position, value):
    """``calc_pointsource_contrib`` -- ``Sources/point_source.py:49-94``.

    Returns ``(dofnos, sum_k N_k(x_s) . value)`` for the cell containing x_s.
    """
    position = np.asarray(position, np.float64)
    value = np.asarray(value)
    c = locate_cell(space.coords, space.cells, position)
    vals, _ = eval_basis(space.order, space.coords[space.cells[c]], position[None])
    return space.cell_dofs[c].copy(), vals[0] @ value


def filament_contrib(space, endpoints, I, no_pts):
    """``FillamentCurrentSource.get_contribution`` -- ``Sources/fillament_current_source.py:62-82``."""
    start, end = np.asarray(endpoints[0], float), np.asarray(endpoints[1], float)
    delta = end - start
    length = np.linalg.norm(delta)
    vec = delta / length * I
    pts = [start + delta * 0.5] if no_pts == 1 else start + delta * np.linspace(0, 1, no_pts)[:, None]
    mag = vec * length / no_pts
    acc = {}
    for pt in pts:
        dn, v = point_source_contrib(space, pt, mag)
        for a, b in zip(dn, v):
            acc[a] = acc.get(a, 0.) + b
    keys = np.array(sorted(acc), np.int64)
    return keys, np.array([acc[k] for k in keys])


def driven_lhs(mats, frequency):
    """``DrivenProblemABC.get_LHS_matrix`` -- ``EMDrivenProblem.py:67-72``."""
    k0 = 2 * np.pi * frequency / c0
    return mats["S"] - k0 ** 2 * mats["M"] + 1j * k0 * mats["S_0"]


def driven_rhs(ndofs, dofnos, contribs, frequency):
    """``DrivenProblemABC.get_RHS`` -- ``EMDrivenProblem.py:74-80``."""
    k0 = 2 * np.pi * frequency / c0
    b = np.zeros(ndofs, np.complex128)
    np.add.at(b, np.asarray(dofnos, np.int64), -1j * k0 * Z0 * np.asarray(contribs))
    return b


def solve_direct(A, b):
    """Direct sparse solve (the reference's UMFPACK path, ``LinalgSolvers.py:262-272``)."""
    return spla.spsolve(A.tocsc(), b)


# ---------------------------------------------------------------------------
# Eigen (resonant cavity, PEC walls)
# ---------------------------------------------------------------------------
def cavity_eigs(space, nev, sigma=1.1, eps_r=None, inv_mu_r=None, dense_limit=4000,
                facet_flags=None):
    """Lowest ``nev`` nonzero k^2 of the PEC-eliminated pencil (S_II, T_II).

    Follows ``EMVectorWaveEigenproblem.py:61-105`` with the spectrum-shift
    correction of ``examples/resonant_cavity/driver.py:66`` applied, i.e. it
    returns lambda = k^2 directly.  Small problems use dense ``eigh``; larger ones
    use shift-invert Lanczos about ``sigma`` selecting nu = 1/(lambda - sigma) > 0
    (largest algebraic), which excludes the gradient null space (nu = -1/sigma).
    """
    mats = assemble_fast(space, eps_r, inv_mu_r, abc=False)
    free = np.flatnonzero(~space.pec_mask(facet_flags))
    S = mats["S"][free][:, free]
    T = mats["M"][free][:, free]
    if len(free) <= dense_limit:
        import scipy.linalg
        w, v = scipy.linalg.eigh(S.toarray(), T.toarray())
        keep = np.flatnonzero(w > sigma)
        return w[keep[:nev]], v[:, keep[:nev]], free
    w, v = spla.eigsh(S.tocsc(), k=nev, M=T.tocsc(), sigma=sigma, which="LA")
    order = np.argsort(w)
    return w[order], v[:, order], free


def analytic_cavity_k2(dims, count, steps=6):
    """Analytic k^2 = (m pi/a)^2 + (n pi/b)^2 + (p pi/c)^2 with TE/TM multiplicity.

    ``functional_tests/test_EigenProblems.py:42-65,111-122``: modes with at most
    one zero index exist; all-nonzero index triples count twice (TE and TM).
    """
    a, b, c = dims
    vals = []
    for m, n, p in itertools.product(range(steps), repeat=3):
        z = (m, n, p).count(0)
        if z < 2:
            k2 = (m * np.pi / a) ** 2 + (n * np.pi / b) ** 2 + (p * np.pi / c) ** 2
            vals.append(k2)
            if z == 0:
                vals.append(k2)
    return np.sort(vals)[:count]


# ---------------------------------------------------------------------------
# Discrete gradient (pins S.G = 0)
# ---------------------------------------------------------------------------
def discrete_gradient(space):
    """Matrix G mapping Lagrange P_k DOFs to Nedelec DOFs with N(G phi) = grad phi.

    Order 1: vertex DOFs, G[e, v_hi] = 1, G[e, v_lo] = -1.
    Order 2: P2 DOFs = vertices then edge midpoints; entries are the nodal dual
    of grad phi (tangential evaluations), computed per cell on the reference.
    """
    edges = space.topo["edges"]
    nv = len(space.coords)
    if space.order == 1:
        ne = len(edges)
        r = np.repeat(np.arange(ne), 2)
        c = edges.ravel()
        v = np.tile([-1.0, 1.0], ne)
        return sp.csr_matrix((v, (r, c)), shape=(ne, nv))
    # order 2: P2 Lagrange basis on the reference tet in barycentrics
    nm = len(MONO)
    lam = [np.zeros(nm) for _ in range(4)]
    lam[0][0] = 1; lam[0][1] = -1; lam[0][2] = -1; lam[0][3] = -1
    lam[1][1] = 1; lam[2][2] = 1; lam[3][3] = 1

    def mul(p, q):
        out = np.zeros(nm)
        for i, mi in enumerate(MONO):
            for j, mj in enumerate(MONO):
                if p[i] != 0 and q[j] != 0:
                    mm = tuple(np.add(mi, mj))
                    out[_MIDX[mm]] += p[i] * q[j]
        return out
    p2 = [mul(lam[i], 2 * lam[i] - np.eye(nm)[0]) for i in range(4)]
    p2 += [4 * mul(lam[a], lam[b]) for a, b in LOCAL_EDGES]
    grads = np.array([np.stack([_deriv(f, 0), _deriv(f, 1), _deriv(f, 2)]) for f in p2])
    Gloc = apply_duals(2, grads).T                    # (20 nedelec, 10 lagrange)
    ne = len(edges)
    nc = len(space.cells)
    ldofs = np.concatenate([space.cells, nv + space.topo["cell_edges"]], axis=1)   # (nc, 10)
    rows = np.repeat(space.cell_dofs, 10, axis=1).ravel()
    cols = np.tile(ldofs, (1, 20)).ravel()
    vals = np.tile(Gloc.ravel(), nc)
    G = sp.coo_matrix((vals, (rows, cols)), shape=(space.ndofs, nv + ne)).tocsr()
    # every cell writes the same value for a shared entry; undo the summation
    cnt = sp.coo_matrix((np.ones(len(vals)), (rows, cols)), shape=(space.ndofs, nv + ne)).tocsr()
    G.data = G.data / cnt.data
    return G


# ---------------------------------------------------------------------------
# Krylov (CPU baseline of the driven solve)
# ---------------------------------------------------------------------------
def cocg_jacobi(A, b, tol=1e-8, maxiter=10000):
    """Jacobi-preconditioned COCG for complex-symmetric A in FP64 (scipy CSR SpMV).

    The CPU counterpart of the device driven solve; the reference's own
    iterative path is ``solve_sparse_system`` (Utilities/LinalgSolvers.py:308-322).
    Returns ``(x, iterations, relative residual history)``.
    """
    A = A.tocsr()
    D = 1.0 / A.diagonal()
    nb = np.linalg.norm(b)
    x = np.zeros_like(b, dtype=np.complex128)
    r = np.asarray(b, np.complex128).copy()
    z = D * r
    p = z.copy()
    rho = r @ z
    hist = []
    for it in range(maxiter):
        q = A @ p
        alpha = rho / (p @ q)
        x += alpha * p
        r -= alpha * q
        z = D * r
        rho_new = r @ z
        p = z + (rho_new / rho) * p
        rho = rho_new
        hist.append(np.linalg.norm(r) / nb)
        if hist[-1] <= tol:
            break
    return x, it + 1, np.array(hist)


def gradient_diag(A, G):
    """diag(G^T A G), column by column: sum_i G_ij (A G)_ij."""
    AG = (A @ G).tocsr()
    return np.asarray(G.multiply(AG).sum(axis=0)).ravel()


def cocg_gradient(A, b, G, tol=1e-8, maxiter=10000, gdiag=None, theta_g=2.0, A_prec=None):
    """COCG with the additive gradient-space preconditioner, FP64 (scipy CSR).

    M^-1 r = D^-1 r + theta_g G diag(G^T A G)^-1 G^T r, D = diag(A); Lagrange nodes whose
    diagonal is zero (none on a mesh without PEC walls) are dropped. Same
    recurrences as ``cocg_jacobi`` with z = M^-1 r.  ``A_prec``: the matrix the diagonals
    are taken from (default A; the device builds them from a complex-shifted driven
    operator).  Returns ``(x, iterations, history)``.
    """
    A = A.tocsr()
    P = A if A_prec is None else A_prec.tocsr()
    if gdiag is None:
        gdiag = gradient_diag(P, G)
    G = G.tocsr().astype(np.complex128)        # complex once, not per product
    Gt = G.T.tocsr()
    D = 1.0 / P.diagonal()
    W = np.zeros(len(gdiag), np.complex128)
    nz = gdiag != 0
    W[nz] = theta_g / gdiag[nz]

    def prec(r):
        return D * r + G @ (W * (Gt @ r))

    nb = np.linalg.norm(b)
    x = np.zeros_like(b, dtype=np.complex128)
    r = np.asarray(b, np.complex128).copy()
    z = prec(r)
    p = z.copy()
    rho = r @ z
    hist = []
    for it in range(maxiter):
        q = A @ p
        alpha = rho / (p @ q)
        x += alpha * p
        r -= alpha * q
        z = prec(r)
        rho_new = r @ z
        p = z + (rho_new / rho) * p
        rho = rho_new
        hist.append(np.linalg.norm(r) / nb)
        if hist[-1] <= tol:
            break
    return x, it + 1, np.array(hist)


def nodal_interpolation(space):
    """Pi: nodal values of a continuous piecewise-linear vector field (3 per vertex,
    column 3 v + d) -> Nedelec DOFs, Pi[i, 3v + d] = lambda_v(x_i) (t_i)_d for the DOF i
    with dual point x_i and physical tangent t_i = J t_ref (the FIAT nodal dual of
    ``_dual_functionals``).  Not a reference function: the auxiliary space of the
    device's Hiptmair-Xu preconditioner.  Pinned by reproduction: for a linear field
    (order 2) or a field a + b x x (order 1), Pi applied to its vertex values is its
    Nedelec interpolant, whose basis expansion equals the field everywhere.
    """
    duals = _dual_functionals(space.order)
    cells = np.asarray(space.cells)
    X = space.coords[cells]
    J = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)
    rows, cols, vals = [], [], []
    for i, (pt, t) in enumerate(duals):
        lam = np.array([1 - pt.sum(), pt[0], pt[1], pt[2]])
        tp = J @ t
        for v in range(4):
            if abs(lam[v]) < 1e-14:
                continue
            for d in range(3):
                rows.append(space.cell_dofs[:, i])
                cols.append(3 * cells[:, v] + d)
                vals.append(lam[v] * tp[:, d])
    rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    shape = (space.ndofs, 3 * len(space.coords))
    Pi = sp.coo_matrix((vals, (rows, cols)), shape=shape).tocsr()
    cnt = sp.coo_matrix((np.ones(len(vals)), (rows, cols)), shape=shape).tocsr()
    Pi.data = Pi.data / cnt.data          # every cell writes the same entry
    return Pi


def cocg_auxiliary(A, b, G, Pi, tol=1e-8, maxiter=10000, gdiag=None, vdiag=None,
                   theta_g=3.0, A_prec=None, theta_v=2.0):
    """COCG with the auxiliary-space preconditioner, FP64 (scipy CSR):
    M^-1 r = D^-1 r + theta_g G diag(G^T A G)^-1 G^T r + theta_v Pi diag(Pi^T A Pi)^-1 Pi^T r
    (zero diagonals dropped; ``gdiag`` / ``vdiag`` precomputed diagonals; the device uses
    theta_g = 3 at order 2 and 2 at order 1, theta_v = 2 for driven systems and 1 for the
    real shifted cavity operator).  ``A_prec``: the matrix D and the two
    auxiliary diagonals are taken from (default A).  Returns ``(x, iterations, history)``."""
    A = A.tocsr()
    P = A if A_prec is None else A_prec.tocsr()
    D = 1.0 / P.diagonal()

    def inv(d):
        w = np.zeros(len(d), np.complex128)
        nz = d != 0
        w[nz] = 1.0 / d[nz]
        return w
    W = theta_g * inv(gradient_diag(P, G) if gdiag is None else gdiag)
    V = theta_v * inv(gradient_diag(P, Pi) if vdiag is None else vdiag)
    G = G.tocsr().astype(np.complex128)
    Gt = G.T.tocsr()
    Pi = Pi.tocsr().astype(np.complex128)
    Pit = Pi.T.tocsr()

    def prec(r):
        return D * r + G @ (W * (Gt @ r)) + Pi @ (V * (Pit @ r))

    nb = np.linalg.norm(b)
    x = np.zeros_like(b, dtype=np.complex128)
    r = np.asarray(b, np.complex128).copy()
    z = prec(r)
    p = z.copy()
    rho = r @ z
    hist = []
    for it in range(maxiter):
        q = A @ p
        alpha = rho / (p @ q)
        x += alpha * p
        r -= alpha * q
        z = prec(r)
        rho_new = r @ z
        p = z + (rho_new / rho) * p
        rho = rho_new
        hist.append(np.linalg.norm(r) / nb)
        if hist[-1] <= tol:
            break
    return x, it + 1, np.array(hist)


# ---------------------------------------------------------------------------
# Surface post-processing: power flux and near-to-far-field transform
# ---------------------------------------------------------------------------
def boundary_face_samples(space, x, n=6):
    """Per boundary face: physical quadrature points, weights (area-scaled), the
    outward unit normal, and E, curl E of the DOF vector ``x`` at the points.

    Yields ``(pts (q,3), w (q,), nrm (3,), E (q,3), curlE (q,3))``.  The rule is the
    collapsed Gauss rule with ``n``^2 points (exact to degree 2n-2 in each direction).
    """
    tp, tw = tri_rule(n)
    x = np.asarray(x)
    for c, lf in zip(space.boundary_face_cell, space.boundary_face_local):
        X = space.coords[space.cells[c]]
        p, q, r = (X[i] for i in LOCAL_FACES[lf])
        nrm = np.cross(q - p, r - p)
        area2 = np.linalg.norm(nrm)
        nrm = nrm / area2
        if nrm @ (X[lf] - p) > 0:            # local face lf is opposite vertex lf
            nrm = -nrm
        pts = p + np.outer(tp[:, 0], q - p) + np.outer(tp[:, 1], r - p)
        vals, curls = eval_basis(space.order, X, pts)
        xl = x[space.cell_dofs[c]]
        yield pts, tw * area2, nrm, np.einsum("pid,i->pd", vals, xl), np.einsum("pid,i->pd", curls, xl)


def surface_flux(space, x, k0, mu_r=1.0, n=6):
    """Real power flux through the mesh boundary, ``PostProcessing/power_flux.py:52-63``:
    (1/(k0 Z0)) int n . (E_r x (-curl E_i / mu_r) + E_i x (curl E_r / mu_r)) dS."""
    total = 0.0
    for _, w, nrm, E, cE in boundary_face_samples(space, x, n):
        f = np.cross(E.real, -cE.imag / mu_r) + np.cross(E.imag, cE.real / mu_r)
        total += w @ (f @ nrm)
    return total / k0 / Z0


def ntff_directions(theta, phi):
    st, ct, sp_, cp = np.sin(theta), np.cos(theta), np.sin(phi), np.cos(phi)
    r_hat = np.array([st * cp, st * sp_, ct])
    theta_hat = np.array([ct * cp, ct * sp_, -st])
    phi_hat = np.array([-sp_, cp, 0.0])
    return r_hat, theta_hat, phi_hat


def surface_ntff(space, x, frequency, theta_deg, phi_deg, n=6):
    """Far field (E_theta, E_phi) at unit radius from the equivalent surface currents,
    ``PostProcessing/surface_ntff.py:36-155`` and ``ntff_expressions.py:22-37``.

    J = n x H with H = j curl E / (k0 Z0), M = -n x E, both weighted by
    exp(j k0 r'.r_hat); E_theta = -f (L_phi + Z0 N_theta), E_phi = f (L_theta - Z0 N_phi),
    f = j k0 exp(-j k0) / (4 pi).  Returns complex (len(theta), 2).
    """
    k0 = frequency * 2 * np.pi / c0
    samples = list(boundary_face_samples(space, x, n))
    out = []
    for th, ph in zip(np.deg2rad(np.atleast_1d(theta_deg)), np.deg2rad(np.atleast_1d(phi_deg))):
        r_hat, theta_hat, phi_hat = ntff_directions(th, ph)
        Nv = np.zeros(3, complex)
        Lv = np.zeros(3, complex)
        for pts, w, nrm, E, cE in samples:
            e = np.exp(1j * k0 * (pts @ r_hat))
            H = 1j * cE / (k0 * Z0)
            Nv += (w * e) @ np.cross(nrm, H)
            Lv += (w * e) @ (-np.cross(nrm, E))
        f = 1j * k0 * np.exp(-1j * k0) / (4 * np.pi)
        out.append((-f * (Lv @ phi_hat + Z0 * (Nv @ theta_hat)),
                    f * (Lv @ theta_hat - Z0 * (Nv @ phi_hat))))
    return np.array(out)


def boundary_edge_cells(space):
    """Cells with at least one edge on the boundary, ``Geometry.py:76-104``."""
    on_b = np.zeros(len(space.topo["edges"]), bool)
    for c, lf in zip(space.boundary_face_cell, space.boundary_face_local):
        fv = set(LOCAL_FACES[lf])
        for le, (a, b) in enumerate(LOCAL_EDGES):
            if a in fv and b in fv:
                on_b[space.topo["cell_edges"][c, le]] = True
    return np.flatnonzero(on_b[space.topo["cell_edges"]].any(axis=1))


def interpolate_boundary(space, f):
    """DOF values of ``f`` (points (q,3) -> complex (q,3)) on the boundary DOFs, zero
    elsewhere: the ``DirichletBC(V, expr, on_boundary).apply`` of
    ``Interpolation/Interpolation.py:71-83``.  Each DOF is its nodal functional,
    f(x_i) . (J t_i), with the dual point and tangent mapped from the reference."""
    duals = _dual_functionals(space.order)
    P = np.array([p for p, _ in duals])
    Tt = np.array([t for _, t in duals])
    mask = space.pec_mask()
    g = np.zeros(space.ndofs, np.complex128)
    for c in np.unique(space.boundary_face_cell):
        X = space.coords[space.cells[c]]
        J = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], axis=1)
        pts = X[0] + P @ J.T
        vals = np.einsum("qd,qd->q", f(pts), Tt @ J.T)
        d = space.cell_dofs[c]
        m = mask[d]
        g[d[m]] = vals[m]
    return g


def em_functional(space, E, g, k0, cells):
    """sum over ``cells`` of g_c^T (S_e - k0^2 T_e) E_c with eps_r = mu_r = 1,
    ``PostProcessing/PostProcessing.py:107-172`` (bilinear, no conjugation)."""
    total = 0.0 + 0.0j
    for c in cells:
        T, S = element_matrices(space.order, space.coords[space.cells[c]])
        d = space.cell_dofs[c]
        total += g[d] @ ((S - k0 ** 2 * T) @ E[d])
    return total


def variational_flux(space, x, k0):
    """``power_flux.VariationalSurfaceFlux.calc_flux`` (``power_flux.py:65-133``):
    conj(F(E, g)) with g = j conj(E on the boundary DOFs) / (k0 Z0), over the
    cells that touch the boundary through an edge."""
    x = np.asarray(x, np.complex128)
    g = 1j * np.where(space.pec_mask(), x, 0).conjugate() / k0 / Z0
    return em_functional(space, x, g, k0, boundary_edge_cells(space)).conjugate()


def ntff_testing_field(rhat, ahat, k0):
    """(j/k0) rhat x (ahat x rhat) exp(j k0 rhat . x), ``variational_ntff.py:31-61``."""
    cf = np.cross(rhat, np.cross(ahat, rhat)) / k0
    return lambda pts: 1j * cf[None, :] * np.exp(1j * k0 * (pts @ rhat))[:, None]


def variational_ntff(space, x, frequency, theta_deg, phi_deg, n=6):
    """``variational_ntff.NTFF.calc_pt`` (``variational_ntff.py:64-131``): the H
    contribution from F(E, g_a) with the boundary interpolant g_a of the testing
    field, the E contribution from the surface L potential."""
    k0 = frequency * 2 * np.pi / c0
    x = np.asarray(x, np.complex128)
    cells = boundary_edge_cells(space)
    samples = list(boundary_face_samples(space, x, n))
    out = []
    for th, ph in zip(np.deg2rad(np.atleast_1d(theta_deg)), np.deg2rad(np.atleast_1d(phi_deg))):
        r_hat, theta_hat, phi_hat = ntff_directions(th, ph)
        EH = [em_functional(space, x, interpolate_boundary(space, ntff_testing_field(r_hat, a, k0)),
                            k0, cells) for a in (theta_hat, phi_hat)]
        Lv = np.zeros(3, complex)
        for pts, w, nrm, E, _ in samples:
            Lv += (w * np.exp(1j * k0 * (pts @ r_hat))) @ (-np.cross(nrm, E))
        f = 1j * k0 * np.exp(-1j * k0) / (4 * np.pi)
        out.append((f * (-(Lv @ phi_hat) + EH[0]), f * (Lv @ theta_hat + EH[1])))
    return np.array(out)
