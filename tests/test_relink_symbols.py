"""The relinked `fem` program (oracle/_ref/fem_gpu, built by
`make -C oracle fem-gpu` where the reference sources are) on the CPU: the
symbol binding the relink promises, and a loud failure without a GPU."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref")
FEM_GPU = os.path.join(REF, "fem_gpu")
DROPIN = ["femEqsCGwJ", "femEqsBiCCSwJ", "femEqsCGwSSOR", "femMatVecMultBig", "femVecVecMultBig",
          "femVecNormBig", "femMatNormBig", "femVecLinCombBig", "femMatSetZeroBig", "femVecSetZeroBig"]

pytestmark = pytest.mark.skipif(not os.path.exists(FEM_GPU) or shutil.which("nm") is None,
                                reason="oracle/_ref/fem_gpu is built where the reference sources are")


def symbols(args):
    out = subprocess.run(["nm"] + args + [FEM_GPU], capture_output=True, text=True, check=True).stdout
    table = {}
    for ln in out.splitlines():
        f = ln.split()
        table[f[-1]] = f[-2]
    return table


def test_binding_of_the_relinked_program():
    s = symbols([])
    # the binding's fem_fill_K is the strong definition; uFEM's survives as fem_fill_K_cpu
    assert s["fem_fill_K"] == "T" and s["fem_fill_K_cpu"] == "T"
    assert s["femEqsCGwJ"] == "U" and s["femMatSetZeroBig"] == "U"
    for n in DROPIN:
        assert s.get(n) in (None, "U"), n     # not defined here: calls bind to libufem_gpu.so
        assert s[n + "_cpu"] == "T", n        # uFEM's CPU function kept
    d = symbols(["-D"])
    # the library's CPU fallback finds the solvers and kernels at run time
    for n in DROPIN[:8]:
        assert d.get(n + "_cpu") == "T", n
    ldd = subprocess.run(["ldd", FEM_GPU], capture_output=True, text=True).stdout
    assert "libufem_gpu.so" in ldd and "not found" not in ldd


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible here")
def test_fails_loudly_without_gpu(tmp_path):
    from ufem_amd.mesh import cantilever, write_fem_input
    fin = str(tmp_path / "in.fem")
    write_fem_input(cantilever(4, 1, 1), fin)
    p = subprocess.run([FEM_GPU, "-p", "-i", fin, "-o", str(tmp_path / "out.res")], capture_output=True,
                       text=True, timeout=60)
    log = p.stdout + p.stderr
    assert p.returncode != 0 and "stiffness matrix on the GPU" in log and "Solution failed" in log
    assert not os.path.exists(tmp_path / "out.res")
