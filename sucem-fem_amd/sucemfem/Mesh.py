"""Tetrahedral meshes and the Nedelec function space (device-resident).

Stands in for the DOLFIN objects the reference passes around: ``dolfin.Mesh`` /
``dolfin.UnitCube`` (Utilities/MeshGenerators.py:33, Testing/Meshes.py:36-47),
``dolfin.MeshFunction`` (Materials.py:135, essential.py:93) and
``dolfin.FunctionSpace(mesh, "Nedelec 1st kind H(curl)", order)``
(ProblemConfigurations/EMProblem.py:100-105).  Numbering, DOF map, sparsity
pattern and PEC DOFs are computed by libsucem_hip on the device.
"""
import ctypes
import xml.etree.ElementTree as ET

import numpy as np

from . import _hip
from .runtime import Handle, get_context


def _tag(el):
    return el.tag.rsplit('}', 1)[-1]


def _read_dolfin_xml_mesh(filename):
    coords = cells = None
    for ev, el in ET.iterparse(filename, events=("start", "end")):
        t = _tag(el)
        if ev == "start":
            if t == "mesh" and el.get("celltype", "tetrahedron") != "tetrahedron":
                raise NotImplementedError("only tetrahedral meshes are supported")
            if t == "vertices":
                coords = np.zeros((int(el.get("size")), 3))
            elif t == "cells":
                cells = np.zeros((int(el.get("size")), 4), np.int64)
            continue
        if t == "vertex":
            coords[int(el.get("index"))] = (float(el.get("x")), float(el.get("y")),
                                            float(el.get("z", 0.0)))
            el.clear()
        elif t == "tetrahedron":
            cells[int(el.get("index"))] = [int(el.get("v%d" % k)) for k in range(4)]
            el.clear()
    if coords is None or cells is None:
        raise ValueError("%s holds no DOLFIN tetrahedral mesh" % filename)
    return coords, cells


def _read_dolfin_xml_meshfunction(filename):
    """(value_type, dim, values) of a DOLFIN XML MeshFunction."""
    vals = None
    vtype, dim = "uint", None
    for ev, el in ET.iterparse(filename, events=("start", "end")):
        t = _tag(el)
        if ev == "start":
            if t == "meshfunction":
                vtype = el.get("type", "uint")
                dim = int(el.get("dim"))
                vals = np.zeros(int(el.get("size")), np.float64 if vtype == "double" else np.int64)
            continue
        if t == "entity":
            v = el.get("value")
            vals[int(el.get("index"))] = float(v) if vtype == "double" else int(v)
            el.clear()
    if vals is None:
        raise ValueError("%s holds no DOLFIN MeshFunction" % filename)
    return vtype, dim, vals


class Mesh(object):
    """Tetrahedral mesh; cell vertex lists are kept sorted (DOLFIN ordering).

    ``Mesh(filename)`` reads a DOLFIN XML mesh (as ``dolfin.Mesh(filename)``,
    e.g. ``examples/albani_bernardi74/driver.py:55``)."""

    def __init__(self, coordinates=None, cells=None):
        if isinstance(coordinates, str):
            coordinates, cells = _read_dolfin_xml_mesh(coordinates)
        self._nfacets = None
        self._coords = None if coordinates is None else np.array(coordinates, np.float64).reshape(-1, 3)
        self._cells = None if cells is None else np.sort(np.array(cells, np.int64).reshape(-1, 4), axis=1)
        self._device = None          # (ctx, Handle) of a device copy
        self._device_coords = None   # host coordinates the device copy was made from

    # -- DOLFIN-like accessors ------------------------------------------------------
    def coordinates(self):
        if self._coords is None:
            self._export()
        return self._coords

    def cells(self):
        if self._cells is None:
            self._export()
        return self._cells

    def num_vertices(self):
        return len(self.coordinates()) if self._coords is not None else int(self._info()[0])

    def num_cells(self):
        return len(self.cells()) if self._cells is not None else int(self._info()[1])

    def topology_dim(self):
        return 3

    def num_facets(self):
        """Number of triangular faces (DOLFIN facet numbering: first encounter)."""
        if self._nfacets is None:
            self._nfacets = int(FunctionSpace(self, "Nedelec 1st kind H(curl)", 1).info()[3])
        return self._nfacets

    # -- device copy ---------------------------------------------------------------
    def _info(self):
        nv, nc = ctypes.c_int64(), ctypes.c_int64()
        _hip.call("sucem_mesh_info", self._device[1].h, ctypes.byref(nv), ctypes.byref(nc))
        return nv.value, nc.value

    def _export(self):
        nv, nc = self._info()
        xyz = np.empty((nv, 3), np.float64)
        cells = np.empty((nc, 4), np.int64)
        _hip.call("sucem_mesh_export", self._device[1].h, _hip.ptr(xyz), _hip.ptr(cells))
        self._coords, self._cells = xyz, cells
        self._device_coords = xyz.copy()

    def device_handle(self, ctx=None):
        """The device mesh, (re)uploaded when the host coordinates changed."""
        ctx = ctx or get_context()
        if self._device is not None and self._device[0] is ctx:
            if self._coords is None or (self._device_coords is not None
                                        and np.array_equal(self._coords, self._device_coords)):
                return self._device[1]
        h = Handle(ctx, "sucem_mesh_destroy")
        xyz = np.ascontiguousarray(self.coordinates(), np.float64)
        cells = np.ascontiguousarray(self.cells(), np.int64)
        _hip.call("sucem_mesh_create", ctx.handle, _hip.ptr(xyz), len(xyz), _hip.ptr(cells),
                  len(cells), ctypes.byref(h.h))
        self._device = (ctx, h)
        self._device_coords = xyz.copy()
        return h

    @classmethod
    def box_on_device(cls, n, scale=(1.0, 1.0, 1.0), shift=(0.0, 0.0, 0.0), ctx=None,
                      export=True):
        """DOLFIN ``UnitCube(nx, ny, nz)`` generated on the device, then scaled and shifted."""
        ctx = ctx or get_context()
        nx, ny, nz = (int(n),) * 3 if np.isscalar(n) else (int(v) for v in n)
        m = cls()
        h = Handle(ctx, "sucem_mesh_destroy")
        sc = _hip.f64(scale, (3,)) * np.ones(3)
        sh = _hip.f64(shift, (3,)) * np.ones(3)
        _hip.call("sucem_mesh_create_box", ctx.handle, nx, ny, nz, _hip.ptr(sc), _hip.ptr(sh),
                  ctypes.byref(h.h))
        m._device = (ctx, h)
        if export:
            m._export()
        return m


def UnitCube(nx, ny, nz):
    """DOLFIN 1.0 ``UnitCube``: Kuhn split of nx*ny*nz boxes into 6 tetrahedra each."""
    return Mesh.box_on_device((nx, ny, nz))


class MeshFunction(object):
    """``dolfin.MeshFunction``: one value per cell (dim 3) or facet (dim 2, in the
    first-encounter facet numbering).  ``MeshFunction(type, mesh, filename)`` reads a
    DOLFIN XML mesh function (``examples/albani_bernardi74/driver.py:57``)."""

    def __init__(self, value_type, mesh, dim):
        self.mesh = mesh
        if isinstance(dim, str):
            _, self.dim, values = _read_dolfin_xml_meshfunction(dim)
            want = mesh.num_cells() if self.dim == 3 else mesh.num_facets()
            if self.dim not in (2, 3) or len(values) != want:
                raise ValueError("mesh function of dim %d with %d values does not fit the mesh"
                                 % (self.dim, len(values)))
            self._values = values
            return
        self.dim = dim
        if dim == 3:
            n = mesh.num_cells()
        elif dim == 2:
            n = mesh.num_facets()
        else:
            raise NotImplementedError("mesh functions on cells (3) and facets (2) only")
        self._values = np.zeros(n, np.float64 if value_type == "double" else np.int64)

    def set_all(self, value):
        self._values[:] = value

    def array(self):
        return self._values

    def __getitem__(self, i):
        return self._values[i]

    def __setitem__(self, i, v):
        self._values[i] = v

    def __len__(self):
        return len(self._values)


class DofMap(object):
    def __init__(self, space):
        self._space = space

    def global_dimension(self):
        return int(self._space.info()[0])

    def max_cell_dimension(self):
        return int(self._space.info()[5])

    def cell_dofs(self, cell_index):
        return self._space.cell_dofs()[cell_index]


class FunctionSpace(object):
    """Nedelec first-kind H(curl) space of order 1 or 2 on a mesh (device-resident)."""

    def __init__(self, mesh, family="Nedelec 1st kind H(curl)", order=1, ctx=None):
        if family not in ("Nedelec 1st kind H(curl)", "N1curl"):
            raise NotImplementedError("only 'Nedelec 1st kind H(curl)' is supported")
        self._mesh = mesh
        self.order = int(order)
        self.ctx = ctx or get_context()
        mh = mesh.device_handle(self.ctx)
        self._mesh_handle = mh
        self.handle = Handle(self.ctx, "sucem_space_destroy")
        _hip.call("sucem_space_create", self.ctx.handle, mh.h, self.order, ctypes.byref(self.handle.h))
        self._info = None
        self._cell_dofs = None
        self._pattern = None
        self._pec = None

    def mesh(self):
        return self._mesh

    def dofmap(self):
        return DofMap(self)

    def info(self):
        if self._info is None:
            info = np.zeros(8, np.int64)
            _hip.call("sucem_space_info", self.handle.h, _hip.ptr(info))
            self._info = info
        return self._info

    def dim(self):
        return int(self.info()[0])

    def cell_dofs(self):
        if self._cell_dofs is None:
            nc = self._mesh.num_cells()
            d = np.empty((nc, int(self.info()[5])), np.int64)
            _hip.call("sucem_space_export_dofmap", self.handle.h, _hip.ptr(d))
            self._cell_dofs = d
        return self._cell_dofs

    def pattern(self):
        """(indptr int64[N+1], indices int32[nnz]) of the CSR sparsity pattern."""
        if self._pattern is None:
            n, nnz = int(self.info()[0]), int(self.info()[1])
            indptr = np.empty(n + 1, np.int64)
            indices = np.empty(nnz, np.int32)
            _hip.call("sucem_space_export_pattern", self.handle.h, _hip.ptr(indptr), _hip.ptr(indices))
            self._pattern = (indptr, indices)
        return self._pattern

    def entities(self):
        info = self.info()
        edges = np.empty((int(info[2]), 2), np.int64)
        faces = np.empty((int(info[3]), 3), np.int64)
        _hip.call("sucem_space_export_entities", self.handle.h, _hip.ptr(edges), _hip.ptr(faces))
        return edges, faces

    def pec_mask(self):
        """Boolean mask of DOFs on boundary facets (PEC-constrained DOFs)."""
        if self._pec is None:
            m = np.empty(self.dim(), np.uint8)
            _hip.call("sucem_space_export_pec", self.handle.h, _hip.ptr(m))
            self._pec = m.astype(bool)
        return self._pec

    def mark_pec(self, facet_flags=None, union=False):
        """PEC DOFs = closure of the flagged facets (None: every boundary facet);
        ``union`` adds to the current set instead of replacing it."""
        flags = None
        if facet_flags is not None:
            flags = np.ascontiguousarray(np.asarray(facet_flags) != 0, np.uint8)
            if len(flags) != int(self.info()[3]):
                raise ValueError("expected %d facet flags, got %d" % (int(self.info()[3]), len(flags)))
        _hip.call("sucem_space_mark_pec", self.ctx.handle, self.handle.h, _hip.ptr(flags),
                  int(bool(union)))
        self._pec = None
        self._info = None


def boundary_facets(mesh):
    """Indices of the facets on the mesh boundary (facets of exactly one cell), in the
    first-encounter facet numbering; marks what ``dolfin.DomainBoundary`` marks."""
    V = FunctionSpace(mesh, "Nedelec 1st kind H(curl)", 1)
    _, faces = V.entities()
    cells = np.sort(mesh.cells(), axis=1)
    loc = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
    tri = np.concatenate([cells[:, list(l)] for l in loc])
    nv = np.int64(mesh.num_vertices())
    key = lambda t: (t[:, 0] * nv + t[:, 1]) * nv + t[:, 2]          # noqa: E731
    uk, cnt = np.unique(key(tri), return_counts=True)
    fk = key(faces)
    return np.flatnonzero(cnt[np.searchsorted(uk, fk)] == 1)
