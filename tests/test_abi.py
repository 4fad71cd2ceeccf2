"""The C-ABI library loads and exports every entry point include/fuss_gpu.h
declares (no GPU work), and the bindings agree with the header's constants."""
import ctypes
import os
import re

import pytest

from fuss_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fuss_gpu.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fuss_[A-Za-z_]+)\s*\(", text)))


def defines():
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"#define\s+(FUSS_\w+)\s+(-?\d+)", open(HEADER).read())}


def test_library_exports_every_declared_symbol():
    if not os.path.isfile(gpu.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(gpu.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_constants_match_header():
    d = defines()
    assert d["FUSS_FP64"] == gpu.FUSS_FP64 and d["FUSS_FP32"] == gpu.FUSS_FP32
    assert d["FUSS_CONS"] == gpu.FUSS_CONS and d["FUSS_PRIM"] == gpu.FUSS_PRIM
    assert d["FUSS_IRS_JACOBI2"] == gpu.FUSS_IRS_JACOBI2 and d["FUSS_IRS_TRIDIAG"] == gpu.FUSS_IRS_TRIDIAG
    assert d["FUSS_ESTATE"] == gpu.FUSS_ESTATE


def test_fortran_binding_constants_match_header():
    src = open(os.path.join(ROOT, "fuss_amd", "fortran", "fuss_gpu_m.f90")).read()
    d = defines()
    for name, val in re.findall(r"(FUSS_\w+)\s*=\s*(\d+)", src):
        assert d[name] == int(val), name
    # every C entry the Fortran module binds exists in the header
    bound = set(re.findall(r"bind\(c, name='(fuss_\w+)'\)", src))
    assert bound <= set(declared())


def test_step_structs_layout():
    # fuss_step_params / fuss_step_result as C lays them out (x86-64 SysV)
    assert ctypes.sizeof(gpu.StepParams) == 48
    assert ctypes.sizeof(gpu.StepResult) == 48
