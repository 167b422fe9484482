"""Host-side mesh input: DOLFIN XML meshes and mesh functions, .femmesh files
(Utilities/MeshIO.py:27-117, Utilities/MeshConverters.py:23-65).  The files are
written here, in the layout of the reference's own mesh files, from the committed
fixtures of those files."""
import numpy as np


def write_dolfin_xml(path, coords, cells):
    with open(path, "w") as f:
        f.write('<?xml version="1.0" encoding="UTF-8"?>\n\n')
        f.write('<dolfin xmlns:dolfin="http://www.fenicsproject.org">\n')
        f.write('  <mesh celltype="tetrahedron" dim="3">\n')
        f.write('    <vertices size="%d">\n' % len(coords))
        for i, p in enumerate(coords):
            f.write('      <vertex index="%d" x="%r" y="%r" z="%r"/>\n' % (i, float(p[0]), float(p[1]), float(p[2])))
        f.write('    </vertices>\n    <cells size="%d">\n' % len(cells))
        for i, c in enumerate(cells):
            f.write('      <tetrahedron index="%d" v0="%d" v1="%d" v2="%d" v3="%d"/>\n'
                    % ((i,) + tuple(int(v) for v in c)))
        f.write('    </cells>\n  </mesh>\n</dolfin>\n')


def write_dolfin_meshfunction(path, dim, values):
    with open(path, "w") as f:
        f.write('<?xml version="1.0" encoding="UTF-8"?>\n')
        f.write('<dolfin xmlns:dolfin="http://www.fenics.org/dolfin/">\n')
        f.write('  <meshfunction type="uint" dim="%d" size="%d">\n' % (dim, len(values)))
        for i, v in enumerate(values):
            f.write('    <entity index="%d" value="%d"/>\n' % (i, v))
        f.write('  </meshfunction>\n</dolfin>\n')


def write_femmesh(path, nodes, tets, props):
    with open(path, "w") as f:
        f.write("BLOCK nodes\n%d\n" % len(nodes))
        for i, p in enumerate(nodes):
            f.write("%d %r %r %r\n" % (i + 1, float(p[0]), float(p[1]), float(p[2])))
        f.write("ENDBLOCK\nBLOCK tris\n1\n1 0 1 2 3\nENDBLOCK\n")
        f.write("BLOCK tets\n%d\n" % len(tets))
        for i, (t, pr) in enumerate(zip(tets, props)):
            f.write("%d %d %d %d %d %d\n" % ((i + 1, pr) + tuple(int(v) + 1 for v in t[::-1])))
        f.write("ENDBLOCK\n")


def test_dolfin_xml_mesh_and_cell_function(tmp_path, golden):
    from sucemfem.Mesh import Mesh, MeshFunction
    d = golden("mesh_albani_bernardi74.npz")
    write_dolfin_xml(str(tmp_path / "m.xml"), d["coords"], d["cells"])
    write_dolfin_meshfunction(str(tmp_path / "m_physical_region.xml"), 3, d["physical_region"])
    m = Mesh(str(tmp_path / "m.xml"))
    assert np.array_equal(m.coordinates(), d["coords"])
    assert np.array_equal(m.cells(), np.sort(d["cells"], axis=1))
    mf = MeshFunction("uint", m, str(tmp_path / "m_physical_region.xml"))
    assert mf.dim == 3 and np.array_equal(mf.array(), d["physical_region"])


def test_meshfunction_size_is_checked(tmp_path, golden):
    import pytest
    from sucemfem.Mesh import Mesh, MeshFunction
    d = golden("mesh_albani_bernardi74.npz")
    write_dolfin_meshfunction(str(tmp_path / "bad.xml"), 3, d["physical_region"][:-1])
    with pytest.raises(ValueError):
        MeshFunction("uint", Mesh(d["coords"], d["cells"]), str(tmp_path / "bad.xml"))


def test_femmesh_reader(tmp_path, golden):
    from sucemfem.Utilities import MeshIO, MeshConverters
    d = golden("mesh_sphere_r1m_4.npz")
    props = np.arange(len(d["tets"])) % 3
    write_femmesh(str(tmp_path / "s.femmesh"), d["nodes"], d["tets"], props)
    r = MeshIO.FemmeshReader(str(tmp_path / "s.femmesh"))
    r.read_meshfile()
    assert np.array_equal(r.get_nodes(), d["nodes"])
    assert np.array_equal(r.get_tet_nodes(), np.sort(d["tets"], axis=1))
    assert np.array_equal(r.get_tet_property_nos(), props)
    assert r.get_mesh_filename() == "s.femmesh"
    m = MeshIO.femmesh_2_dolfin_mesh(str(tmp_path / "s.femmesh"))
    assert np.array_equal(m.cells(), np.sort(d["tets"], axis=1))
    lm = MeshConverters.dolfin_mesh_2_listmesh(m)
    assert lm["ElementNodes"].dtype == np.int32 and np.array_equal(lm["Nodes"], d["nodes"])


def test_femmesh_accessors_need_read(tmp_path):
    import pytest
    from sucemfem.Utilities import MeshIO
    r = MeshIO.FemmeshReader(str(tmp_path / "none.femmesh"))
    with pytest.raises(MeshIO.UninitialisedException):
        r.get_nodes()


def _boundary_triangles(cells):
    loc = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
    tri = np.sort(np.concatenate([np.sort(cells, axis=1)[:, list(l)] for l in loc]), axis=1)
    u, cnt = np.unique(tri, axis=0, return_counts=True)
    return u[cnt == 1]


def test_gmsh_reader_rectangular_prism(golden):
    """functional_tests/data/meshes/rectangular_prism.msh (gmsh 2.2, the 1 x 0.5 x 0.25 prism
    of rectangular_prism.geo with the walls in physical group 1), read by the conversion rules
    of sandbox/pec_labels/pec_labels.py:27-443: tetrahedra make the mesh, the tagged triangles
    are exactly its boundary facets."""
    import os
    from sucemfem.Utilities.GmshIO import GmshMesh
    g = GmshMesh(os.path.join(os.path.dirname(__file__), "golden", "rectangular_prism.msh"))
    xyz, cells = g.coordinates_and_cells()
    assert xyz.shape == (107, 3) and cells.shape == (329, 4)
    # vertices numbered by ascending gmsh node id
    used = np.unique(g.elements[4][1])
    ref = dict(zip(g.node_ids.tolist(), g.node_xyz))
    assert np.array_equal(xyz, np.array([ref[i] for i in used]))
    assert np.all(np.diff(cells, axis=1) > 0)
    X = xyz[cells]
    J = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)
    vol = np.abs(np.linalg.det(J)).sum() / 6
    assert abs(vol - 0.125) < 1e-12
    assert np.allclose(xyz.min(axis=0), 0) and np.allclose(xyz.max(axis=0), [1, 0.5, 0.25])
    assert g.cell_regions() is None                     # every tetrahedron in region 0
    tri, tag = g.tagged_triangles()
    assert len(tri) == 156 and set(tag.tolist()) == {1}
    bnd = _boundary_triangles(cells)
    assert np.array_equal(np.unique(np.sort(tri, axis=1), axis=0), bnd)


def test_gmsh_reader_errors(tmp_path):
    import pytest
    from sucemfem.Utilities.GmshIO import GmshMesh, GmshFormatError
    p = tmp_path / "bad.msh"
    p.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n2\n1 0 0 0\n$EndNodes\n")
    with pytest.raises(GmshFormatError):
        GmshMesh(str(p))
    p.write_text("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n0\n$EndNodes\n$Elements\n0\n"
                 "$EndElements\n")
    with pytest.raises(GmshFormatError):
        GmshMesh(str(p))
    p.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n4\n1 0 0 0\n2 1 0 0\n3 0 1 0\n"
                 "4 0 0 1\n$EndNodes\n$Elements\n1\n1 4 2 0 1 1 2 3 9\n$EndElements\n")
    with pytest.raises(GmshFormatError):
        GmshMesh(str(p)).coordinates_and_cells()


def test_femmesh_malformed_block(tmp_path):
    import pytest
    from sucemfem.Utilities import MeshIO
    p = tmp_path / "bad.femmesh"
    p.write_text("BLOCK nodes\n2\n1 0 0 0\n3 1 0 0\nENDBLOCK\n")
    with pytest.raises(AssertionError):
        MeshIO.FemmeshReader(str(p)).read_meshfile()
    p.write_text("BLOCK nodes\n2\n1 0 0 0\nENDBLOCK\n")
    with pytest.raises(AssertionError):
        MeshIO.FemmeshReader(str(p)).read_meshfile()
