"""Convert the reference's own test fixtures into small ``.npz`` files.

Run once in a container where the reference tree is mounted::

    python tests/golden/make_fixtures.py /root/reference

The reference stores its golden data as Python-2 pickles.  They are decoded
here by :class:`InertPickleReader`, an opcode-level parser that walks the
pickle byte stream and rebuilds only plain containers and numpy arrays.  It
never imports or calls anything named in the file: ``GLOBAL`` opcodes become
inert ``("global", module, name)`` tokens and ``REDUCE``/``BUILD`` become inert
records that a small whitelist (numpy ``ndarray`` reconstruction, ``dtype`` and
``scalar``) turns into data afterwards.  Anything outside that whitelist is an
error.

Fixtures written (data only: inputs and expected outputs):

* ``driven_order1_inscribed_tet.npz`` -- ``LHS_matrix.npy`` (24x24 complex)
  and ``RHS_vector.pickle`` (24 complex) of
  ``sucemfem/ProblemConfigurations/tests/test_EMDrivenProblem.py:36-82``.
* ``filament_order2.npz`` -- ``sucemfem/Sources/tests/fillament_source_test_data.pickle``
  (``test_fillament_source.py:48-82``).
* ``driven_order2_dipole_x.npz`` -- ``reference_dofs-2-0.149896229-0.0499654096667.pickle``
  (order-2 driven solve, 5530 complex DOFs; generator
  ``PostProcessing/tests/data/ntff_generate_dofs.py:40-95``).
* ``driven_order2_filament_x.npz`` -- ``power_flux_reference_dofs_...pickle`` (1314 DOFs,
  generator ``PostProcessing/tests/data/power_flux_generate_dofs.py:26-102``) plus the
  cells and the (6-digit) vertex coordinates of its DOLFIN mesh XML, which
  ``test_power_flux.py:42`` loads.
* ``driven_order2_filament9_x.npz`` -- ``examples/filament_source/data/dofs-2-...-100.pickle``
  (30186 DOFs).
* ``ntff_surface_order2.npz``, ``ntff_variational_order2.npz`` -- far fields of the
  solution in ``driven_order2_dipole_x.npz`` at six angles
  (``PostProcessing/tests/test_ntff.py:58-103``).
* ``ntff_interpolant.npz`` -- ``interpolant_test_data.pickle`` (``test_ntff.py:36-55``).
* ``mesh_rectangular_prism.npz`` -- ``functional_tests/data/meshes/rectangular_prism.xml`` and
  its facet regions (PEC walls = 1; ``functional_tests/test_EigenProblems.py:69-84``).
* ``mesh_albani_bernardi74.npz`` -- ``examples/albani_bernardi74/mesh/*fig2VII*.xml`` with
  the cell regions (1000: eps_r 16, 1001: eps_r 1; ``examples/albani_bernardi74/driver.py:53-61``).
* ``mesh_sphere_r1m_4.npz`` -- ``sandbox/solvers/meshes/sphere-r1m-4.femmesh`` (nodes, tets,
  property numbers), the input format of ``Utilities/MeshIO.py:39-117``.
* ``unitcube1_topology.npz`` -- the boundary-edge count of UnitCube(1,1,1)
  asserted by ``sucemfem/tests/test_Geometry.py:31-35``.
"""
import os
import re
import sys
import codecs
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class InertPickleReader(object):
    """Parse a protocol 0-2 pickle without executing anything from it."""

    def __init__(self, data):
        self.data = data
        self.pos = 0

    # -- low level readers -------------------------------------------------
    def _line(self):
        end = self.data.index(b"\n", self.pos)
        s = self.data[self.pos:end]
        self.pos = end + 1
        return s

    def _take(self, n):
        s = self.data[self.pos:self.pos + n]
        if len(s) != n:
            raise ValueError("truncated pickle")
        self.pos += n
        return s

    def load(self):
        stack, memo, marks = [], {}, []

        def pop_mark():
            m = marks.pop()
            items = stack[m:]
            del stack[m:]
            return items

        while True:
            op = self._take(1)
            if op == b"(":                      # MARK
                marks.append(len(stack))
            elif op == b".":                    # STOP
                return stack.pop()
            elif op == b"\x80":                 # PROTO
                self._take(1)
            elif op == b"N":
                stack.append(None)
            elif op == b"\x88":
                stack.append(True)
            elif op == b"\x89":
                stack.append(False)
            elif op == b"I":                    # INT (text)
                s = self._line()
                if s == b"00":
                    stack.append(False)
                elif s == b"01":
                    stack.append(True)
                else:
                    stack.append(int(s))
            elif op == b"L":                    # LONG (text)
                stack.append(int(self._line().rstrip(b"L")))
            elif op == b"F":                    # FLOAT (text)
                stack.append(float(self._line()))
            elif op == b"J":                    # BININT
                stack.append(struct.unpack("<i", self._take(4))[0])
            elif op == b"K":                    # BININT1
                stack.append(self._take(1)[0])
            elif op == b"M":                    # BININT2
                stack.append(struct.unpack("<H", self._take(2))[0])
            elif op == b"\x8a":                 # LONG1
                n = self._take(1)[0]
                stack.append(int.from_bytes(self._take(n), "little", signed=True))
            elif op == b"G":                    # BINFLOAT
                stack.append(struct.unpack(">d", self._take(8))[0])
            elif op == b"S":                    # STRING (quoted repr, escapes)
                s = self._line().strip()
                q = s[:1]
                if q not in (b"'", b'"') or s[-1:] != q:
                    raise ValueError("bad STRING opcode")
                stack.append(codecs.escape_decode(s[1:-1])[0])
            elif op == b"T":                    # BINSTRING
                n = struct.unpack("<i", self._take(4))[0]
                stack.append(self._take(n))
            elif op == b"U":                    # SHORT_BINSTRING
                n = self._take(1)[0]
                stack.append(self._take(n))
            elif op == b"V":                    # UNICODE (raw-unicode-escape)
                stack.append(self._line().decode("raw_unicode_escape"))
            elif op == b"X":                    # BINUNICODE
                n = struct.unpack("<I", self._take(4))[0]
                stack.append(self._take(n).decode("utf-8"))
            elif op == b"c":                    # GLOBAL -> inert token
                mod = self._line().decode("ascii")
                name = self._line().decode("ascii")
                stack.append(("global", mod, name))
            elif op == b"p":                    # PUT
                memo[int(self._line())] = stack[-1]
            elif op == b"q":                    # BINPUT
                memo[self._take(1)[0]] = stack[-1]
            elif op == b"r":                    # LONG_BINPUT
                memo[struct.unpack("<I", self._take(4))[0]] = stack[-1]
            elif op == b"g":                    # GET
                stack.append(memo[int(self._line())])
            elif op == b"h":                    # BINGET
                stack.append(memo[self._take(1)[0]])
            elif op == b"j":                    # LONG_BINGET
                stack.append(memo[struct.unpack("<I", self._take(4))[0]])
            elif op == b"t":                    # TUPLE
                stack.append(tuple(pop_mark()))
            elif op == b")":
                stack.append(())
            elif op == b"\x85":
                stack.append((stack.pop(),))
            elif op == b"\x86":
                b_ = stack.pop(); a_ = stack.pop()
                stack.append((a_, b_))
            elif op == b"\x87":
                c_ = stack.pop(); b_ = stack.pop(); a_ = stack.pop()
                stack.append((a_, b_, c_))
            elif op == b"l":                    # LIST
                stack.append(list(pop_mark()))
            elif op == b"]":
                stack.append([])
            elif op == b"a":                    # APPEND
                v = stack.pop()
                stack[-1].append(v)
            elif op == b"e":                    # APPENDS
                items = pop_mark()
                stack[-1].extend(items)
            elif op == b"d":                    # DICT
                items = pop_mark()
                stack.append(dict(zip(items[0::2], items[1::2])))
            elif op == b"}":
                stack.append({})
            elif op == b"s":                    # SETITEM
                v = stack.pop(); k = stack.pop()
                stack[-1][k] = v
            elif op == b"u":                    # SETITEMS
                items = pop_mark()
                for k, v in zip(items[0::2], items[1::2]):
                    stack[-1][k] = v
            elif op == b"R":                    # REDUCE -> inert record
                args = stack.pop(); fn = stack.pop()
                stack.append({"__call__": fn, "args": args, "state": None})
            elif op == b"\x81":                 # NEWOBJ -> inert record
                args = stack.pop(); cls = stack.pop()
                stack.append({"__call__": cls, "args": args, "state": None})
            elif op == b"b":                    # BUILD -> attach state
                state = stack.pop()
                obj = stack[-1]
                if not (isinstance(obj, dict) and "__call__" in obj):
                    raise ValueError("BUILD on a non-record")
                obj["state"] = state
            else:
                raise ValueError("unsupported pickle opcode %r at %d" % (op, self.pos - 1))


def _as_bytes(s):
    if isinstance(s, bytes):
        return s
    if isinstance(s, str):                      # py2 str decoded as latin-1
        return s.encode("latin-1")
    raise TypeError(type(s))


def materialise(obj):
    """Turn the inert records of numpy reconstruction into numpy data."""
    if isinstance(obj, list):
        return [materialise(v) for v in obj]
    if isinstance(obj, tuple) and not (len(obj) == 3 and obj[0] == "global"):
        return tuple(materialise(v) for v in obj)
    if isinstance(obj, dict) and "__call__" not in obj:
        return {materialise(k): materialise(v) for k, v in obj.items()}
    if isinstance(obj, dict):
        fn = obj["__call__"]
        if fn == ("global", "numpy", "dtype"):
            code = _as_bytes(obj["args"][0]).decode("ascii")
            dt = np.dtype(code)
            st = obj["state"]
            if st is not None and len(st) > 1:
                order = _as_bytes(st[1]).decode("ascii") if not isinstance(st[1], str) else st[1]
                if order in "<>":
                    dt = dt.newbyteorder(order)
            return dt
        if fn == ("global", "numpy.core.multiarray", "_reconstruct"):
            st = obj["state"]
            _, shape, dtype, fortran, raw = st[:5]
            dtype = materialise(dtype)
            if isinstance(raw, list):
                raise ValueError("object arrays are not supported")
            arr = np.frombuffer(_as_bytes(raw), dtype=dtype).copy()
            arr = arr.reshape(shape, order="F" if fortran else "C")
            return arr.astype(dtype.newbyteorder("="))
        if fn == ("global", "numpy.core.multiarray", "scalar"):
            dtype = materialise(obj["args"][0])
            raw = _as_bytes(obj["args"][1])
            return np.frombuffer(raw, dtype=dtype)[0]
        raise ValueError("refusing unknown callable %r" % (fn,))
    if isinstance(obj, bytes):
        try:
            return obj.decode("ascii")
        except UnicodeDecodeError:
            return obj
    return obj


def read_pickle(path):
    with open(path, "rb") as f:
        return materialise(InertPickleReader(f.read()).load())


def read_dolfin_mesh_cells(path):
    """Cells (sorted vertex ids) written by DOLFIN into a mesh XML file."""
    txt = open(path).read()
    cells = re.findall(r'<tetrahedron index="(\d+)" v0="(\d+)" v1="(\d+)" v2="(\d+)" v3="(\d+)"', txt)
    cells = np.array([[int(v) for v in c[1:]] for c in sorted(cells, key=lambda c: int(c[0]))],
                     dtype=np.int64)
    nv = int(re.search(r'<vertices size="(\d+)"', txt).group(1))
    return nv, cells


def read_dolfin_mesh_vertices(path):
    """Vertex coordinates as written (6 significant digits) in a DOLFIN mesh XML file."""
    txt = open(path).read()
    v = re.findall(r'<vertex index="(\d+)" x="([^"]+)" y="([^"]+)" z="([^"]+)"', txt)
    return np.array([[float(c) for c in r[1:]] for r in sorted(v, key=lambda r: int(r[0]))],
                    dtype=np.float64)


def read_dolfin_meshfunction(path):
    """Values (by entity index) of a DOLFIN XML MeshFunction."""
    txt = open(path).read()
    ent = re.findall(r'<entity index="(\d+)" value="(-?\d+)"', txt)
    vals = np.zeros(int(re.search(r'size="(\d+)"', txt).group(1)), np.int64)
    for i, v in ent:
        vals[int(i)] = int(v)
    return vals


def read_femmesh(path):
    """Nodes, tets (file order, 0-based node ids) and property numbers of a .femmesh file."""
    lines = open(path).read().splitlines()
    nodes = tets = props = None
    i = 0
    while i < len(lines):
        if lines[i].startswith("BLOCK nodes"):
            n = int(lines[i + 1].split()[0])
            nodes = np.zeros((n, 3))
            for l in lines[i + 2:i + 2 + n]:
                sp = l.split()
                nodes[int(sp[0]) - 1] = [float(v) for v in sp[1:4]]
            i += 2 + n
        elif lines[i].startswith("BLOCK tets"):
            n = int(lines[i + 1].split()[0])
            tets = np.zeros((n, 4), np.int64)
            props = np.zeros(n, np.int64)
            for l in lines[i + 2:i + 2 + n]:
                sp = l.split()
                tets[int(sp[0]) - 1] = [int(v) - 1 for v in sp[2:6]]
                props[int(sp[0]) - 1] = int(sp[1])
            i += 2 + n
        else:
            i += 1
    return nodes, tets, props


def main(ref):
    out = HERE
    # 1. order-1 driven LHS/RHS on InscribedTet
    lhs = np.load(os.path.join(ref, "sucemfem/ProblemConfigurations/tests/LHS_matrix.npy"),
                  allow_pickle=False)
    rhs = read_pickle(os.path.join(ref, "sucemfem/ProblemConfigurations/tests/RHS_vector.pickle"))
    np.savez_compressed(os.path.join(out, "driven_order1_inscribed_tet.npz"),
                        lhs=np.asarray(lhs, np.complex128), rhs=np.asarray(rhs, np.complex128),
                        frequency=1e8, source_coord=np.array([0., 0., 0.]),
                        source_value=np.array([2., -1., 3.]))
    # 2. order-2 filament source contributions
    d = read_pickle(os.path.join(ref, "sucemfem/Sources/tests/fillament_source_test_data.pickle"))
    np.savez_compressed(os.path.join(out, "filament_order2.npz"),
                        order=int(d["order"]), I=float(d["I"]),
                        source_endpoints=np.asarray(d["source_endpoints"], np.float64),
                        domain_size=np.asarray(d["domain_size"], np.float64),
                        max_edge_len=float(d["max_edge_len"]),
                        no_integration_points=int(d["no_integration_points"]),
                        dofnos=np.asarray(d["dofnos"], np.int64),
                        rhs_contribs=np.asarray(d["rhs_contribs"], np.float64))
    # 3. order-2 driven solution, point source at the origin
    d = read_pickle(os.path.join(ref, "sucemfem/PostProcessing/tests/data/"
                                 "reference_dofs-2-0.149896229-0.0499654096667.pickle"))
    np.savez_compressed(os.path.join(out, "driven_order2_dipole_x.npz"),
                        x=np.asarray(d["x"], np.complex128), order=int(d["order"]),
                        domain_size=np.asarray(d["domain_size"], np.float64),
                        max_edge_len=float(d["max_edge_len"]), freq=float(d["freq"]))
    # 4. order-2 driven solution, lambda/10 filament, with the DOLFIN mesh cells
    base = ("sucemfem/PostProcessing/tests/data/power_flux_reference_dofs_f-1000000000.000000"
            "_o-2_s-0.074948_l-0.100000_h-0.166667")
    d = read_pickle(os.path.join(ref, base + ".pickle"))
    nv, cells = read_dolfin_mesh_cells(os.path.join(ref, base + "_mesh.xml"))
    keep = {k: v for k, v in d.items() if isinstance(v, (np.ndarray, float, int, np.floating))}
    coords = read_dolfin_mesh_vertices(os.path.join(ref, base + "_mesh.xml"))
    np.savez_compressed(os.path.join(out, "driven_order2_filament_x.npz"),
                        mesh_nv=nv, mesh_cells=cells, mesh_coords=coords,
                        **{k: np.asarray(v) for k, v in keep.items()})
    # 5. larger order-2 driven solution (UnitCube(9), lambda/4 filament)
    d = read_pickle(os.path.join(ref, "examples/filament_source/data/"
                                 "dofs-2-0.299792458-0.0499654096667-0.0749481145-100.pickle"))
    keep = {k: v for k, v in d.items() if isinstance(v, (np.ndarray, float, int, np.floating))}
    np.savez_compressed(os.path.join(out, "driven_order2_filament9_x.npz"),
                        **{k: np.asarray(v) for k, v in keep.items()})
    # 6. far fields of the solution of 3. (surface and variational NTFF) at six angles
    for kind in ("surface", "variational"):
        d = read_pickle(os.path.join(ref, "sucemfem/PostProcessing/tests/data/reference_%s_ntff"
                                     "-2-0.149896229-0.0499654096667.pickle" % kind))
        ff = d["ff_result_data"]
        np.savez_compressed(os.path.join(out, "ntff_%s_order2.npz" % kind),
                            theta_deg=np.asarray(ff["theta_deg"], np.float64),
                            phi_deg=np.asarray(ff["phi_deg"], np.float64),
                            E_ff=np.asarray(ff["E_ff"], np.complex128),
                            E_theta_analytical=np.asarray(ff["E_theta_analytical"], np.complex128))
    # 7. variational NTFF testing-function interpolant (test_ntff.py:36-55)
    d = read_pickle(os.path.join(ref, "sucemfem/PostProcessing/tests/data/interpolant_test_data.pickle"))
    np.savez_compressed(os.path.join(out, "ntff_interpolant.npz"), k0=float(d["k0"]),
                        rhat=np.asarray(d["rhat"], np.float64),
                        ahats=np.asarray(d["ahats"], np.float64),
                        coords=np.asarray(d["coords"], np.float64),
                        vals=np.asarray(d["vals"], np.complex128))
    # 8. meshes read from files by the reference's drivers and functional tests
    base = "functional_tests/data/meshes/rectangular_prism"
    np.savez_compressed(os.path.join(out, "mesh_rectangular_prism.npz"),
                        coords=read_dolfin_mesh_vertices(os.path.join(ref, base + ".xml")),
                        cells=read_dolfin_mesh_cells(os.path.join(ref, base + ".xml"))[1],
                        facet_region=read_dolfin_meshfunction(
                            os.path.join(ref, base + "_facet_region.xml")))
    base = "examples/albani_bernardi74/mesh/albani_bernardi74_fig2VII"
    np.savez_compressed(os.path.join(out, "mesh_albani_bernardi74.npz"),
                        coords=read_dolfin_mesh_vertices(os.path.join(ref, base + ".xml")),
                        cells=read_dolfin_mesh_cells(os.path.join(ref, base + ".xml"))[1],
                        physical_region=read_dolfin_meshfunction(
                            os.path.join(ref, base + "_physical_region.xml")))
    nodes, tets, props = read_femmesh(os.path.join(ref, "sandbox/solvers/meshes/sphere-r1m-4.femmesh"))
    np.savez_compressed(os.path.join(out, "mesh_sphere_r1m_4.npz"), nodes=nodes, tets=tets,
                        property_nos=props)
    # 9. UnitCube(1,1,1) boundary-edge count
    np.savez_compressed(os.path.join(out, "unitcube1_topology.npz"), n_boundary_edges=18)
    for f in sorted(os.listdir(out)):
        if f.endswith(".npz"):
            z = np.load(os.path.join(out, f))
            print(f, {k: (z[k].shape, str(z[k].dtype)) for k in z.files})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
