"""The C-ABI library loads and exports every entry point the header declares (no GPU calls)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sucem_hip.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sucem_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "sucem_solve_driven" in names and "sucem_assemble" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    from sucemfem import _hip
    lib = _hip.load()
    for name in declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _hip.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r" T (sucem_[a-z0-9_]+)", out))
    assert set(declared()) <= exported


def test_python_binding_covers_header():
    from sucemfem import _hip
    assert set(_hip.SIGNATURES) == set(declared())


def test_version_and_error_string_without_gpu():
    from sucemfem import _hip
    lib = _hip.load()
    assert lib.sucem_version() >= 10000
    assert isinstance(lib.sucem_last_error(), bytes)


def test_null_handles_are_rejected_without_gpu():
    import ctypes
    from sucemfem import _hip
    lib = _hip.load()
    st = lib.sucem_space_info(None, None)
    assert st == _hip.ERR_INVALID
    assert b"NULL" in lib.sucem_last_error()
    o = _hip.SolveOpts()
    lib.sucem_solve_opts_default(ctypes.byref(o))
    assert o.tol == 1e-8 and o.precision == 64 and o.method == _hip.METHOD_COCG


def test_struct_layouts_match_the_header():
    """The ctypes mirrors of the option and statistics structs have the header's field
    offsets and sizes (gcc on include/sucem_hip.h)."""
    import ctypes
    import shutil
    import subprocess
    import tempfile
    from sucemfem import _hip
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    structs = {"sucem_solve_opts": _hip.SolveOpts, "sucem_solve_stats": _hip.SolveStats,
               "sucem_eig_opts": _hip.EigOpts, "sucem_eig_stats": _hip.EigStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sucem_hip.h"',
             'int main(void) {']
    for cname, py in structs.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append('return 0; }')
    inc = os.path.join(os.path.dirname(__file__), "..", "include")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        with open(src, "w") as fh:
            fh.write("\n".join(lines))
        subprocess.check_call(["gcc", "-I", inc, src, "-o", exe])
        out = subprocess.check_output([exe]).decode().split("\n")
    got = dict(l.rsplit(" ", 1) for l in out if l)
    for cname, py in structs.items():
        assert int(got["%s size" % cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got["%s.%s" % (cname, f)]) == getattr(py, f).offset, (cname, f)
