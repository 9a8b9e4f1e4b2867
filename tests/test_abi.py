"""The C-ABI library loads without a GPU and exports every function that
include/shear_rk3.h declares; the ctypes mirror matches the header's structs."""
import ctypes
import os
import re

import pytest

import shear_nnf_amd as snf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shear_rk3.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(snf_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_python_exports():
    assert sorted(snf.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol():
    if not os.path.exists(snf.LIB_PATH):
        pytest.fail("libshear_rk3.so is not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(snf.LIB_PATH)
    missing = [name for name in _declared() if not hasattr(lib, name)]
    assert not missing, missing


def _struct_fields(name):
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), src, re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        typ, names = decl.split(None, 1)
        for n in names.split(","):
            n = n.strip()
            m = re.match(r"(\w+)\[(\d+)\]", n)
            fields.append((m.group(1), typ, int(m.group(2))) if m else (n, typ, 1))
    return fields


@pytest.mark.parametrize("cname,pycls", [("snf_params", snf.SnfParams),
                                          ("snf_run_status", snf.SnfRunStatus)])
def test_struct_layout_matches_header(cname, pycls):
    ctype = {"int32_t": ctypes.c_int32, "double": ctypes.c_double}
    want = _struct_fields(cname)
    got = [(f[0], f[1]) for f in pycls._fields_]
    assert [w[0] for w in want] == [g[0] for g in got]
    for (n, typ, cnt), (_, pyt) in zip(want, got):
        exp = ctype[typ] if cnt == 1 else ctype[typ] * cnt
        assert ctypes.sizeof(pyt) == ctypes.sizeof(exp), n
