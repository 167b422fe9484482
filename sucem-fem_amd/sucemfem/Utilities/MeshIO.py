"""Mesh file input (sucemfem/Utilities/MeshIO.py:27-117).

The ``.femmesh`` format is a sequence of named blocks::

    BLOCK <name>
    <row count>
    <row>                 one entity per line, its 1-based id first
    ...
    ENDBLOCK

``nodes`` rows are ``id x y z``; ``tets`` rows are ``id property v1 v2 v3 v4`` with
1-based vertex ids; ``tris`` and any other block are skipped.  The whole file is split
into blocks in one scan, each wanted block is converted by a single numpy parse of its
rows, and the rows are placed by their ids, which must be a permutation of
1..count.  Tetrahedron vertex lists are stored sorted (DOLFIN's mesh ordering), with
one property number per tetrahedron.  A malformed block raises
:class:`FemmeshFormatError` (an ``AssertionError``, as the reference's checks are).

``femmesh_2_dolfin_mesh`` returns a :class:`sucemfem.Mesh.Mesh`.  DOLFIN XML meshes and
mesh functions are read by ``Mesh(filename)`` and ``MeshFunction(type, mesh, filename)``;
gmsh ``.msh`` files by :mod:`sucemfem.Utilities.GmshIO`.
"""
from __future__ import division

import os

import numpy as N

from . import MeshConverters


class UninitialisedException(Exception):
    """An accessor of :class:`FemmeshReader` was called before ``read_meshfile()``."""


class FemmeshFormatError(AssertionError):
    """The file does not follow the block layout described in the module docstring."""


def read_femmesh_blocks(path):
    """{block name: list of its row strings} for every block of a ``.femmesh`` file."""
    with open(path) as f:
        lines = f.read().splitlines()
    blocks = {}
    i, n = 0, len(lines)
    while i < n:
        head = lines[i].split()
        if len(head) < 2 or head[0] != 'BLOCK':
            i += 1
            continue
        name = ' '.join(head[1:])
        try:
            count = int(lines[i + 1].split()[0])
        except (IndexError, ValueError):
            raise FemmeshFormatError("%s: block %r has no row count" % (path, name))
        stop = i + 2 + count
        tail = lines[stop].split() if stop < n else []
        if count < 0 or stop > n or not tail or tail[0] != 'ENDBLOCK':
            raise FemmeshFormatError("%s: block %r: expected %d rows and ENDBLOCK"
                                     % (path, name, count))
        blocks[name] = lines[i + 2:stop]
        i = stop + 1
    return blocks


def _rows_to_table(rows, ncols, dtype, what):
    flat = N.array(' '.join(rows).split(), dtype=dtype)
    if flat.size != len(rows) * ncols:
        raise FemmeshFormatError("%s rows must have %d columns" % (what, ncols))
    return flat.reshape(len(rows), ncols)


def _place_by_id(ids, values, what):
    slot = ids.astype(N.int64) - 1
    if not N.array_equal(N.sort(slot), N.arange(len(slot))):
        raise FemmeshFormatError("%s ids are not a permutation of 1..%d" % (what, len(slot)))
    out = N.empty_like(values)
    out[slot] = values
    return out


class FemmeshReader(object):
    """Reader of one ``.femmesh`` file; call :meth:`read_meshfile` before the getters."""

    def __init__(self, mesh_filename):
        self.mesh_filename = mesh_filename
        self.read = False

    def read_meshfile(self):
        blocks = read_femmesh_blocks(self.mesh_filename)
        if 'nodes' in blocks:
            t = _rows_to_table(blocks['nodes'], 4, N.float64, "node")
            self.nodes = _place_by_id(t[:, 0], t[:, 1:4].copy(), "node")
            self.no_nodes = len(self.nodes)
        if 'tets' in blocks:
            t = _rows_to_table(blocks['tets'], 6, N.int64, "tet")
            ids = t[:, 0]
            self.tet_nodes = _place_by_id(ids, N.sort(t[:, 2:6] - 1, axis=1).astype(N.int32),
                                          "tet")
            self.tet_property_nos = _place_by_id(ids, t[:, 1].astype(N.int32), "tet")
            self.no_tets = len(self.tet_nodes)
        self.read = True

    def _need(self, attr):
        if not self.read or not hasattr(self, attr):
            raise UninitialisedException("Call read_meshfile() method of %s first"
                                         % self.__class__)
        return getattr(self, attr)

    def get_mesh_filename(self):
        self._need('read')
        return os.path.basename(self.mesh_filename)

    def get_mesh_dirname(self):
        self._need('read')
        return os.path.dirname(self.mesh_filename)

    def get_tet_nodes(self):
        return self._need('tet_nodes')

    def get_tet_property_nos(self):
        return self._need('tet_property_nos')

    def get_nodes(self):
        return self._need('nodes')


def femmesh_2_dolfin_mesh(femmesh_file):
    """Read a femmesh file into a mesh (``MeshIO.py:114-117``)."""
    return MeshConverters.femmesh_reader_2_dolfin_mesh(FemmeshReader(femmesh_file))
