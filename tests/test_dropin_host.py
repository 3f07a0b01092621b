"""The Fortran binding (pixelflow_amd/fortran/pf_binding.f90) against the C-ABI, on the
CPU: pf_desc and pf_stats declare the same components in the same order and types as
include/pixelflow.h (as mirrored by pixelflow_amd.solver), every interface binds a
symbol the library exports, and the drop-in executables are built and linked to it."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

FSRC = os.path.join(ROOT, "pixelflow_amd", "fortran", "pf_binding.f90")
BIN = os.path.join(ROOT, "pixelflow_amd", "_build", "bin")
FTYPES = {"integer(c_int32_t)": ctypes.c_int32, "real(c_double)": ctypes.c_double}


def _fortran_type(name):
    text = open(FSRC).read()
    body = re.search(rf"type, bind\(C\) :: {name}\n(.*?)end type {name}", text, re.S).group(1)
    comps = []
    for line in body.splitlines():
        line = line.split("!")[0].strip()
        if not line:
            continue
        ftype, names = line.split("::")
        for item in names.split(","):
            nm = item.split("=")[0].strip()
            count = 1
            mm = re.match(r"(\w+)\((\d+)\)", nm)
            if mm:
                nm, count = mm.group(1), int(mm.group(2))
            comps.append((nm, FTYPES[ftype.strip()], count))
    return comps


def _ctypes_layout(struct):
    out = []
    for nm, t in struct._fields_:
        if hasattr(t, "_length_"):
            out.append((nm, t._type_, t._length_))
        else:
            out.append((nm, t, 1))
    return out


@pytest.mark.parametrize("name", ["pf_desc", "pf_stats"])
def test_fortran_types_match_c_layout(name):
    from pixelflow_amd import solver
    struct = {"pf_desc": solver.PfDesc, "pf_stats": solver.PfStats}[name]
    assert _fortran_type(name) == _ctypes_layout(struct)


def test_fortran_interfaces_bind_exported_symbols():
    from pixelflow_amd import solver
    names = set(re.findall(r'bind\(C, name="(pf_\w+)"\)', open(FSRC).read()))
    assert names and names <= set(solver.EXPORTS)
    L = ctypes.CDLL(solver.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n


@pytest.mark.parametrize("exe", ["ibm_2d_uniform", "ibm_2d_drag", "ibm_2d_backstep", "ibm_3d_uniform",
                                 "ibm_3d_air_condition"])
def test_dropin_executables_link_the_library(exe):
    path = os.path.join(BIN, exe)
    if not os.path.exists(path):
        pytest.skip("Fortran executables not built (flang missing)")
    out = subprocess.run(["ldd", path], capture_output=True, text=True, check=True).stdout
    assert "libpixelflow.so" in out and "not found" not in out
