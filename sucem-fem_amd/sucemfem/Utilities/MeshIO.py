"""Mesh file input (sucemfem/Utilities/MeshIO.py:27-117).

``FemmeshReader`` parses the block-structured ``.femmesh`` format (``BLOCK nodes``,
``BLOCK tets``, ``BLOCK tris``; 1-based ids) with the reference's accessors;
``femmesh_2_dolfin_mesh`` returns a :class:`sucemfem.Mesh.Mesh`.  DOLFIN XML meshes
and mesh functions are read by ``Mesh(filename)`` and
``MeshFunction(type, mesh, filename)``.
"""
from __future__ import division

import os
import re

import numpy as N

from . import MeshConverters


class UninitialisedException(Exception):
    pass


def uninit_error(fn):
    def check_fn(self):
        if not self.read:
            raise UninitialisedException(
                "Call read_meshfile() method of %s first" % self.__class__)
        return fn(self)
    return check_fn


class FemmeshReader(object):
    block_start_re = re.compile(r'^BLOCK (.*)')

    def __init__(self, mesh_filename):
        self.mesh_filename = mesh_filename
        self.read = False
        self.block_parse_funs = dict(nodes=self.parse_nodes, tets=self.parse_tets,
                                     tris=self.skip_block)

    def _count(self, lines):
        return int(next(lines).split()[0])

    def parse_nodes(self, lines):
        self.no_nodes = n = self._count(lines)
        self.nodes = N.zeros((n, 3))
        seen = N.zeros(n, bool)
        for _ in range(n):
            sp = next(lines).split()
            i = int(sp[0]) - 1
            assert 0 <= i < n
            self.nodes[i] = [float(x) for x in sp[1:4]]
            seen[i] = True
        assert next(lines).split()[0] == 'ENDBLOCK'
        assert seen.all()

    def parse_tets(self, lines):
        self.no_tets = n = self._count(lines)
        self.tet_nodes = N.zeros((n, 4), dtype=N.int32)
        self.tet_property_nos = N.zeros(n, dtype=N.int32)   # one per tet
        seen = N.zeros(n, bool)
        for _ in range(n):
            sp = next(lines).split()
            i = int(sp[0]) - 1
            assert 0 <= i < n
            self.tet_property_nos[i] = int(sp[1])
            self.tet_nodes[i] = sorted(int(x) - 1 for x in sp[2:6])
            seen[i] = True
        assert next(lines).split()[0] == 'ENDBLOCK'
        assert seen.all()

    def skip_block(self, lines):
        for l in lines:
            if l.split() and l.split()[0] == 'ENDBLOCK':
                return

    def find_block_fun(self, l):
        m = self.block_start_re.match(l)
        if not m:
            return None
        return self.block_parse_funs[m.groups()[0].strip()]

    def read_meshfile(self):
        with open(self.mesh_filename) as fo:
            lines = iter(fo)
            for l in lines:
                bfun = self.find_block_fun(l)
                if bfun:
                    bfun(lines)
        self.read = True

    @uninit_error
    def get_mesh_filename(self):
        return os.path.basename(self.mesh_filename)

    @uninit_error
    def get_mesh_dirname(self):
        return os.path.dirname(self.mesh_filename)

    @uninit_error
    def get_tet_nodes(self):
        return self.tet_nodes

    @uninit_error
    def get_tet_property_nos(self):
        return self.tet_property_nos

    @uninit_error
    def get_nodes(self):
        return self.nodes


def femmesh_2_dolfin_mesh(femmesh_file):
    """Read a femmesh file into a mesh (``MeshIO.py:114-117``)."""
    return MeshConverters.femmesh_reader_2_dolfin_mesh(FemmeshReader(femmesh_file))
