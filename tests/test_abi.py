"""The C ABI: libtbec.so loads without a GPU and exports every function include/tbec.h
declares; the ctypes structs match the header's field lists."""
import ctypes as C
import os
import re

from conftest import ROOT
from trapbec_amd import tbec

HEADER = os.path.join(ROOT, "include", "tbec.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tbec_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(tbec.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(tbec.EXPORTS)


def _struct_fields(name):
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    m = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + ";", src)
    body = m.group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        parts = decl.split(None, 1)
        fields += [f.strip() for f in parts[1].split(",")]
    return fields


def test_struct_layouts_match_header():
    assert _struct_fields("tbec_params") == [f for f, _ in tbec.TbecParams._fields_]
    assert _struct_fields("tbec_diag") == [f for f, _ in tbec.TbecDiag._fields_]
    assert _struct_fields("tbec_solve_result") == [f for f, _ in tbec.TbecSolveResult._fields_]


def test_version_and_error_without_gpu():
    lib = tbec.load()
    assert b"gfx950" in lib.tbec_version()
    h = C.c_void_p()
    p = tbec.TbecParams(0, 8, 8, 64, 0, -1, 1.0, 1.0, 1.0, 0, 0, 0, 0, 0, 0, 1e-3, 0, 0)
    rc = lib.tbec_create(C.byref(p), C.byref(h))
    assert rc == -1 and b"grid" in lib.tbec_last_error()
