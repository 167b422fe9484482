"""Structured cube meshes (sucemfem/Utilities/MeshGenerators.py:26-53)."""
from __future__ import division

import numpy as N

from ..Mesh import Mesh


def get_centred_cube(domain_size, max_edge_len, centred_element_coordinate=None):
    """Cube mesh centred around the origin.

    n = ceil(sqrt(2)*L/h) boxes per axis (DOLFIN UnitCube, generated on the
    device), coordinates scaled by L and shifted by -L/2.  With
    ``centred_element_coordinate`` the mesh is translated so that the point sits at
    the centroid of the containing cell whose midpoint is nearest to it; on ties
    (a point shared by several cells) the lowest-index such cell is used, which on
    the Kuhn split is the first tetrahedron of the box.
    """
    domain_size = N.asarray(domain_size, dtype=N.float64) * N.ones(3)
    n = N.array(N.ceil(N.sqrt(2) * domain_size / max_edge_len), dtype=N.int64)
    mesh = Mesh.box_on_device(tuple(int(v) for v in n), scale=domain_size, shift=-domain_size / 2)
    if centred_element_coordinate is not None:
        from ..Sources.point_source import containing_cells_midpoints
        p = N.asarray(centred_element_coordinate, dtype=N.float64)
        cand, mids = containing_cells_midpoints(mesh, p)
        d = N.linalg.norm(mids - p, axis=1)
        best = N.flatnonzero(d <= d.min() * (1 + 1e-12) + 1e-300)[0]
        centre = mids[best]
        coords = mesh.coordinates()
        coords -= centre
    return mesh
