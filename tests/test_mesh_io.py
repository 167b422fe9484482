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
