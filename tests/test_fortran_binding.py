"""The Fortran bind(C) module include/shear_rk3.f90 (what the reference driver
uses in place of `call cross`) against the C header include/shear_rk3.h:

* it compiles with the image's Fortran compiler (ROCm flang);
* c_sizeof and the field offsets of its snf_params / snf_run_status equal
  sizeof / offsetof of the C structs (a C program compiled with gcc);
* a program that takes the address of every interface procedure links
  against libshear_rk3.so (every bind(C) name resolves) and runs on a host
  without a GPU (it calls no entry point);
* every entry point the header declares has an interface in the module.
"""
import os
import re
import shutil
import subprocess

import pytest

import shear_nnf_amd as snf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOD = os.path.join(ROOT, "include", "shear_rk3.f90")
HEADER = os.path.join(ROOT, "include", "shear_rk3.h")
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"

pytestmark = pytest.mark.skipif(not os.path.exists(FLANG), reason="no Fortran compiler")

PARAMS = ["mgalx", "my", "mgalz", "Re", "alp", "gam", "Ly", "s", "chi", "CFL", "explicit_visc",
          "nhist", "readflag", "uprim", "vbulk", "etime", "dumptime", "dtimag", "precision",
          "device", "ydealias"]
STATUS = ["steps_done", "stop", "time", "Deltat", "xoffb", "xofft", "umax", "CFLv",
          "nhist_records", "stop_reason"]

F_PROBE = """
program probe
  use iso_c_binding
  use snf_gpu
  implicit none
  type(snf_params), target :: p
  type(snf_run_status), target :: st
  integer(c_intptr_t) :: b
  b = transfer(c_loc(p), b)
  write(*, '(a, i0)') 'params ', c_sizeof(p)
%(poff)s
  b = transfer(c_loc(st), b)
  write(*, '(a, i0)') 'status ', c_sizeof(st)
%(soff)s
end program probe
"""

F_LINK = """
program linkall
  use iso_c_binding
  use snf_gpu
  implicit none
  type(c_funptr) :: f(%(n)d)
%(body)s
  write(*, '(a, i0)') 'linked ', size(f)
end program linkall
"""

C_PROBE = """
#include <stddef.h>
#include <stdio.h>
#include "shear_rk3.h"
int main(void) {
  printf("params %%zu\\n", sizeof(snf_params));
%(poff)s
  printf("status %%zu\\n", sizeof(snf_run_status));
%(soff)s
  return 0;
}
"""


def _declared():
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(snf_[a-z0-9_]+)\s*\(", src)))


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, (cmd, r.stdout, r.stderr)
    return r.stdout


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("fortran"))
    _run([FLANG, "-c", MOD, "-module-dir", d, "-o", os.path.join(d, "snf_gpu.o")], d)
    return d


def test_every_entry_point_has_an_interface():
    src = open(MOD).read().lower()
    names = set(re.findall(r'bind\(c, name="(snf_[a-z0-9_]+)"\)', src))
    assert names == set(_declared())


def test_struct_sizes_and_offsets_match_c(built):
    d = built
    poff = "\n".join(f"  write(*, '(a, i0)') 'p.{f} ', transfer(c_loc(p%{f}), b) - b"
                     for f in PARAMS)
    soff = "\n".join(f"  write(*, '(a, i0)') 's.{f} ', transfer(c_loc(st%{f}), b) - b"
                     for f in STATUS)
    with open(os.path.join(d, "probe.f90"), "w") as fh:
        fh.write(F_PROBE % {"poff": poff, "soff": soff})
    _run([FLANG, "probe.f90", "snf_gpu.o", "-I", d, "-o", "probe_f"], d)
    fout = _run([os.path.join(d, "probe_f")], d)
    cp = "\n".join(f'  printf("p.{f} %zu\\n", offsetof(snf_params, {f}));' for f in PARAMS)
    cs = "\n".join(f'  printf("s.{f} %zu\\n", offsetof(snf_run_status, {f}));' for f in STATUS)
    with open(os.path.join(d, "probe.c"), "w") as fh:
        fh.write(C_PROBE % {"poff": cp, "soff": cs})
    _run(["gcc", "-I", os.path.join(ROOT, "include"), "probe.c", "-o", "probe_c"], d)
    cout = _run([os.path.join(d, "probe_c")], d)
    parse = lambda s: dict(line.split() for line in s.strip().splitlines())
    f, c = parse(fout), parse(cout)
    assert f == c
    assert f["params"] == "144" and f["status"] == "88"


def test_module_links_against_the_library(built):
    if not os.path.exists(snf.LIB_PATH):
        pytest.fail("libshear_rk3.so is not built: run __graft_entry__.build()")
    d = built
    names = _declared()
    body = "\n".join(f"  f({q + 1}) = c_funloc({n})" for q, n in enumerate(names))
    with open(os.path.join(d, "linkall.f90"), "w") as fh:
        fh.write(F_LINK % {"n": len(names), "body": body})
    libdir = os.path.dirname(snf.LIB_PATH)
    _run([FLANG, "linkall.f90", "snf_gpu.o", "-I", d, "-L", libdir, "-lshear_rk3",
          f"-Wl,-rpath,{libdir}", "-o", "linkall"], d)
    out = _run([os.path.join(d, "linkall")], d)
    assert out.split() == ["linked", str(len(names))]
