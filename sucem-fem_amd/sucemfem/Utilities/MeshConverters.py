"""Mesh format conversion (sucemfem/Utilities/MeshConverters.py:23-65).

A "listmesh" is a dict with ``Nodes`` (vertex coordinates) and ``ElementNodes``
(tetrahedron vertex ids).  Cell vertex lists are sorted as DOLFIN's ``order()``
sorts them; the cell order is kept.
"""
import numpy as N

from ..Mesh import Mesh


def listmesh_2_dolfin_mesh(listmesh, reorder=False):
    """Mesh from a listmesh (``MeshConverters.py:23-42``); cells are always stored with
    sorted vertices, which is what ``reorder=True`` asks DOLFIN for."""
    return Mesh(N.asarray(listmesh['Nodes'], N.float64),
                N.asarray(listmesh['ElementNodes'], N.int64))


def dolfin_mesh_2_listmesh(dolfin_mesh):
    """``MeshConverters.py:44-56``: Nodes and ElementNodes (int32) copies."""
    return {'Nodes': dolfin_mesh.coordinates().copy(),
            'ElementNodes': N.array(dolfin_mesh.cells(), dtype=N.int32, copy=True)}


def femmesh_reader_2_dolfin_mesh(femmesh_reader, reorder=True):
    """``MeshConverters.py:58-65``."""
    if not femmesh_reader.read:
        femmesh_reader.read_meshfile()
    listmesh = dict(FemmeshFilename=femmesh_reader.get_mesh_filename(),
                    FemmeshDir=femmesh_reader.get_mesh_dirname(),
                    Nodes=femmesh_reader.get_nodes(),
                    ElementNodes=femmesh_reader.get_tet_nodes())
    return listmesh_2_dolfin_mesh(listmesh, reorder=reorder)
