"""Tetrahedral meshes and the Nedelec function space (device-resident).

Stands in for the DOLFIN objects the reference passes around: ``dolfin.Mesh`` /
``dolfin.UnitCube`` (Utilities/MeshGenerators.py:33, Testing/Meshes.py:36-47),
``dolfin.MeshFunction`` (Materials.py:135, essential.py:93) and
``dolfin.FunctionSpace(mesh, "Nedelec 1st kind H(curl)", order)``
(ProblemConfigurations/EMProblem.py:100-105).  Numbering, DOF map, sparsity
pattern and PEC DOFs are computed by libsucem_hip on the device.
"""
import ctypes

import numpy as np

from . import _hip
from .runtime import Handle, get_context


class Mesh(object):
    """Tetrahedral mesh; cell vertex lists are kept sorted (DOLFIN ordering)."""

    def __init__(self, coordinates=None, cells=None):
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
    """Minimal ``dolfin.MeshFunction``: one value per entity of dimension ``dim``."""

    def __init__(self, value_type, mesh, dim):
        self.mesh = mesh
        self.dim = dim
        n = mesh.num_cells() if dim == 3 else 0
        self._values = np.zeros(n, np.int64)

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
