"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: this runs without a GPU)."""

import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ufem_amd", "libufem_gpu.so")
HEADERS = [os.path.join(ROOT, "include", h) for h in ("ufem_fem_math.h", "ufem_gpu.h", "ufem_monte.h")]

DECL = re.compile(r"^\s*(?:const\s+)?(?:int|long|double|void|char|ufem_sys|ufem_model|ufem_net|ufem_batch)"
                  r"\s*\*?\s*(\w+)\s*\(", re.M)


def declared():
    names = []
    for h in HEADERS:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += DECL.findall(txt)
    return sorted(set(names))


def test_headers_declare_the_dropin_symbols():
    names = declared()
    for n in ("femEqsCGwJ", "femEqsBiCCSwJ", "femEqsCGwSSOR", "femMatVecMultBig", "femVecVecMultBig",
              "femVecNormBig", "femMatNormBig", "femVecLinCombBig", "femMatSetZeroBig", "femVecSetZeroBig",
              "ufem_gpu_fill_K", "ufem_gpu_set_mirror_policy", "ufem_gpu_mark_dirty", "ufem_gpu_release",
              "ufem_sys_create",
              "ufem_sys_solve", "ufem_gpu_element_stiffness", "ufem_sys_update",
              "monte_dlib_interface_type", "monte_nums_of_vars", "monte_ivar_name", "monte_ovar_name",
              "monte_init_lib_stuff", "monte_clean_lib_stuff", "monte_solution"):
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libufem_gpu.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_library_loads_without_gpu():
    import ctypes
    L = ctypes.CDLL(LIB)
    for n in declared():
        assert hasattr(L, n)


def test_types_match_reference_layout():
    """tMatrix / tVector keep uFEM's member order and LP64 sizes."""
    from ufem_amd.gpu import tMatrix, tVector
    assert ctypes_sizeof(tMatrix) == 8 * 8
    assert ctypes_sizeof(tVector) == 5 * 8
    assert [f for f, _ in tMatrix._fields_] == ["type", "rows", "cols", "len", "pos", "data", "frompos", "defpos"]


def ctypes_sizeof(t):
    import ctypes
    return ctypes.sizeof(t)
