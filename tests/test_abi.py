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
        fields += [re.sub(r"\[.*\]", "", f).strip() for f in parts[1].split(",")]
    return fields


def test_struct_layouts_match_header():
    assert _struct_fields("tbec_params") == [f for f, _ in tbec.TbecParams._fields_]
    assert _struct_fields("tbec_diag") == [f for f, _ in tbec.TbecDiag._fields_]
    assert _struct_fields("tbec_solve_result") == [f for f, _ in tbec.TbecSolveResult._fields_]
    assert _struct_fields("tbec_ic") == [f for f, _ in tbec.TbecIc._fields_]
    assert _struct_fields("tbec_physics") == [f for f, _ in tbec.TbecPhysics._fields_]
    assert _struct_fields("tbec_reference_ic") == [f for f, _ in tbec.TbecReferenceIc._fields_]
    # sizes as the C compiler lays them out (no implicit padding in any of them)
    assert C.sizeof(tbec.TbecParams) == 6 * 4 + 12 * 8
    assert C.sizeof(tbec.TbecPhysics) == 4 * 4 + 4 * 8
    assert C.sizeof(tbec.TbecReferenceIc) == 4 * 4 + 8
    assert C.sizeof(tbec.TbecIc) == 3 * 8 + 8 * 8 + 16 * 8 + 8 * 4 + 2 * 4


def test_version_and_error_without_gpu():
    lib = tbec.load()
    assert b"gfx950" in lib.tbec_version()
    h = C.c_void_p()
    p = tbec.TbecParams(0, 8, 8, 64, 0, -1, 1.0, 1.0, 1.0, 0, 0, 0, 0, 0, 0, 1e-3, 0, 0)
    rc = lib.tbec_create(C.byref(p), C.byref(h))
    assert rc == -1 and b"grid" in lib.tbec_last_error()


def _params(**kw):
    d = dict(nx=8, ny=8, nz=8, prec=64, sum_mode=0, tile=-1, Lx=1.0, Ly=1.0, Lz=1.0,
             dt_re=1e-3)
    d.update(kw)
    p = tbec.TbecParams()
    for k, v in d.items():
        setattr(p, k, v)
    return p


def test_create_rejects_bad_arguments_before_touching_the_gpu():
    """Every argument check of tbec_create runs before any HIP call, so each rejection
    (return code TBEC_ERR_ARG, a message naming the problem, no handle) is visible on
    a machine without a GPU."""
    lib = tbec.load()
    cases = [(dict(nx=0), b"grid"), (dict(nz=-3), b"grid"), (dict(prec=16), b"prec"),
             (dict(prec=32, sum_mode=3), b"sum_mode"), (dict(sum_mode=2), b"FP64"),
             (dict(Ly=0.0), b"box"), (dict(Lz=float("nan")), b"box"),
             (dict(tile=5), b"tile"), (dict(tile=-2), b"tile"),
             # one x-plane of 12000 x 12000 complex128 (padded) exceeds 2^31 bytes
             (dict(ny=12000, nz=12000), b"x-plane")]
    for kw, word in cases:
        h = C.c_void_p(12345)
        rc = lib.tbec_create(C.byref(_params(**kw)), C.byref(h))
        assert rc == -1, kw
        assert word in lib.tbec_last_error(), (kw, lib.tbec_last_error())
        assert h.value is None, kw
    assert lib.tbec_create(None, None) == -1
