"""gmsh mesh input: the gmsh -> DOLFIN conversion of sandbox/pec_labels/pec_labels.py:27-443
(``gmsh_to_dolfin_mesh``, after DOLFIN's ``meshconvert``).

Reads ASCII gmsh 2.x files: ``$MeshFormat``, ``$Nodes`` (``id x y z`` rows) and
``$Elements`` (``id type ntags tag... node...`` rows); lines starting with ``#`` are
skipped.  Element type 4 (4-node tetrahedron) makes the mesh, type 2 (3-node triangle)
carries facet tags, and other types are ignored, as in the conversion.  The conversion
rules kept here:

* the mesh's vertices are the nodes used by tetrahedra, numbered in ascending node id
  (the order the conversion gets from a set of node ids), cells in element order with
  their vertex lists sorted (DOLFIN mesh ordering);
* the first tag of an element is its physical region; a cell mesh function
  ``physical_region`` exists when some tetrahedron has a non-zero one;
* a triangle with a non-zero physical region marks the mesh facet with the same three
  vertices in a facet mesh function ``facet_region`` (0 elsewhere); a tagged triangle
  that is not a facet of the mesh is an error.  (The sandbox script slices the
  triangles' node list with a stride of 2 per triangle; each triangle's own three nodes
  are used here.)

The facet function is indexed in the mesh's facet numbering, which the device computes
(first encounter), so building it needs a HIP device; reading nodes and elements does not.
"""
from __future__ import division

import numpy as np

TET, TRI = 4, 2


class GmshFormatError(ValueError):
    pass


def _sections(path):
    """{section name: list of its lines} for the $Name ... $EndName sections."""
    out, name, body = {}, None, None
    with open(path) as f:
        for raw in f:
            line = raw.strip()
            if not line or line.startswith('#'):
                continue
            if name is None:
                if line.startswith('$') and not line.startswith('$End'):
                    name, body = line[1:], []
                continue
            if line == '$End' + name:
                out[name] = body
                name = None
            else:
                body.append(line)
    if name is not None:
        raise GmshFormatError("%s: section $%s is not closed" % (path, name))
    return out


class GmshMesh(object):
    """The nodes and elements of one gmsh 2.x ASCII file."""

    def __init__(self, filename):
        self.filename = filename
        sec = _sections(filename)
        for need in ('MeshFormat', 'Nodes', 'Elements'):
            if need not in sec:
                raise GmshFormatError("%s: no $%s section" % (filename, need))
        version = sec['MeshFormat'][0].split()
        if not version[0].startswith('2') or version[1] != '0':
            raise GmshFormatError("%s: only ASCII gmsh 2.x files are supported (format %s)"
                                  % (filename, ' '.join(version)))
        nodes = sec['Nodes']
        count = int(nodes[0])
        if len(nodes) != count + 1:
            raise GmshFormatError("%s: $Nodes announces %d nodes, holds %d"
                                  % (filename, count, len(nodes) - 1))
        table = np.array(' '.join(nodes[1:]).split(), np.float64).reshape(count, 4)
        self.node_ids = table[:, 0].astype(np.int64)
        self.node_xyz = table[:, 1:4].copy()
        elems = sec['Elements']
        count = int(elems[0])
        if len(elems) != count + 1:
            raise GmshFormatError("%s: $Elements announces %d elements, holds %d"
                                  % (filename, count, len(elems) - 1))
        self.elements = {}               # type -> (tags (n, ntags), nodes (n, k))
        rows = {}
        for line in elems[1:]:
            v = [int(t) for t in line.split()]
            etype, ntags = v[1], v[2]
            rows.setdefault(etype, []).append((v[3:3 + ntags], v[3 + ntags:]))
        for etype, rr in rows.items():
            if len(set((len(t), len(n)) for t, n in rr)) != 1:
                raise GmshFormatError("%s: elements of type %d differ in tag or node count"
                                      % (filename, etype))
            self.elements[etype] = (np.array([t for t, _ in rr], np.int64).reshape(len(rr), -1),
                                    np.array([n for _, n in rr], np.int64))
        if TET not in self.elements:
            raise GmshFormatError("%s holds no tetrahedra" % filename)

    def _vertex_numbering(self):
        """Node ids used by tetrahedra (ascending) and their coordinates."""
        used = np.unique(self.elements[TET][1])
        order = np.argsort(self.node_ids, kind='stable')
        sorted_ids = self.node_ids[order]
        pos = np.minimum(np.searchsorted(sorted_ids, used), len(sorted_ids) - 1)
        if np.any(sorted_ids[pos] != used):
            raise GmshFormatError("%s: a tetrahedron uses an undefined node" % self.filename)
        return used, self.node_xyz[order[pos]]

    def coordinates_and_cells(self):
        """(coordinates (nv, 3), cells (nc, 4)) of the tetrahedral mesh."""
        used, xyz = self._vertex_numbering()
        cells = np.searchsorted(used, self.elements[TET][1])
        return xyz, np.sort(cells, axis=1)

    def cell_regions(self):
        """Physical region of each tetrahedron (first tag), or None when all are 0."""
        tags = self.elements[TET][0]
        if tags.shape[1] == 0 or not np.any(tags[:, 0]):
            return None
        return tags[:, 0].copy()

    def tagged_triangles(self):
        """(vertex triples in mesh numbering (n, 3), physical region (n,)) of the triangles
        with a non-zero physical region."""
        if TRI not in self.elements:
            return np.zeros((0, 3), np.int64), np.zeros(0, np.int64)
        tags, nodes = self.elements[TRI]
        if tags.shape[1] == 0:
            return np.zeros((0, 3), np.int64), np.zeros(0, np.int64)
        keep = tags[:, 0] != 0
        used, _ = self._vertex_numbering()
        tri = nodes[keep]
        pos = np.searchsorted(used, tri)
        pos = np.minimum(pos, len(used) - 1)
        if np.any(used[pos] != tri):
            raise GmshFormatError("%s: a tagged triangle uses a node no tetrahedron uses"
                                  % self.filename)
        return np.sort(pos, axis=1), tags[keep, 0].copy()


def gmsh_to_dolfin_mesh(filename):
    """(mesh, physical_region cell function or None, facet_region facet function or None)
    of a gmsh 2.x file."""
    from ..Mesh import Mesh, MeshFunction
    g = GmshMesh(filename)
    coords, cells = g.coordinates_and_cells()
    mesh = Mesh(coords, cells)
    regions = g.cell_regions()
    cell_mf = None
    if regions is not None:
        cell_mf = MeshFunction("uint", mesh, 3)
        cell_mf.array()[:] = regions
    tri, tag = g.tagged_triangles()
    facet_mf = None
    if len(tri):
        from ..Mesh import FunctionSpace
        _, faces = FunctionSpace(mesh, "Nedelec 1st kind H(curl)", 1).entities()
        nv = np.int64(len(coords))
        key = lambda t: (t[:, 0] * nv + t[:, 1]) * nv + t[:, 2]       # noqa: E731
        fk = key(np.sort(np.asarray(faces, np.int64), axis=1))
        order = np.argsort(fk)
        tk = key(tri)
        pos = np.minimum(np.searchsorted(fk[order], tk), len(fk) - 1)
        if np.any(fk[order][pos] != tk):
            raise GmshFormatError("%s: a tagged triangle is not a facet of the mesh" % filename)
        facet_mf = MeshFunction("uint", mesh, 2)
        facet_mf.array()[order[pos]] = tag
    return mesh, cell_mf, facet_mf
