"""uFEM's own `fem` program relinked against libufem_gpu.so, end to end.

oracle/_ref/fem_gpu is the reference's `fem` with fem_fill_K bound to
ufem_amd/integration/fem_fill_K_gpu.c and the solver/"Big" names bound to
the drop-in library (its own CPU functions kept under <name>_cpu); it reads
an input file, assembles K on the GPU, applies loads and supports on the CPU
(fem/fem_sol.c:1119-1247), solves on the GPU and writes the result file
with uFEM's own code (fem/fem_io.c:648-734).  oracle/_ref/fem_ref is the
same program unchanged.  Both are built by `make -C oracle fem-ref fem-gpu`
from the reference sources.

Checks: the GPU route was taken (UFEM_TRACE lines, the fill message), the
result file's integer parts equal the unchanged program's, and its numbers
agree to the 7 digits the file prints (displacements, reactions, element
results; entries below 1e-6 of the largest one compared in absolute terms).
The -fbc case is also held against the committed golden het8_fbc.res."""

import os
import subprocess

import numpy as np
import pytest

from tests.resfile import read_res
from ufem_amd.mesh import cantilever, input_ids, lognormal_E, write_fem_input

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref")
FEM_GPU = os.path.join(REF, "fem_gpu")
FEM_REF = os.path.join(REF, "fem_ref")

needs_build = pytest.mark.skipif(
    not (os.path.exists(FEM_GPU) and os.path.exists(FEM_REF)),
    reason="oracle/_ref/fem_gpu and fem_ref are built by `make -C oracle fem-ref fem-gpu` "
           "where the reference sources are")


def run(prog, args, out, env=None):
    e = dict(os.environ)
    e.update(env or {})
    p = subprocess.run([prog] + args + ["-o", out], capture_output=True, text=True, timeout=120, env=e)
    log = p.stdout + p.stderr
    assert p.returncode == 0 and "Solution done" in log, log[-2000:]
    return log


def agree(a, b, rel=2e-6, floor=1e-6):
    scale = np.abs(b).max() if b.size else 0.0
    return bool(np.all(np.abs(a - b) <= rel * np.abs(b) + floor * scale + 1e-300))


def compare(mine, ref):
    for k in ("header", "n_act", "n_dof_len", "n_nodes", "n_elems", "n_res"):
        assert mine[k] == ref[k], k
    for k in ("node_id", "dof_map", "r_node", "r_id", "r_dof", "elem_id", "res_from"):
        assert np.array_equal(mine[k], ref[k]), k
    for k in ("u", "R", "res"):
        assert agree(mine[k], ref[k]), (k, np.abs(mine[k] - ref[k]).max(), np.abs(ref[k]).max())


CASES = {
    # tag: (mesh, input with IDs, program flags, solver the GPU must run)
    "c1_cg": (lambda: cantilever(20, 4, 4), False, ["-p"], "femEqsCGwJ"),
    "c1_ssor": (lambda: cantilever(20, 4, 4), False, ["-p", "-ssor"], "femEqsCGwSSOR"),
    "c1_bicgs": (lambda: cantilever(20, 4, 4), False, ["-p", "-nes"], "femEqsBiCCSwJ"),
    "c1_ids": (lambda: cantilever(20, 4, 4), True, [], "femEqsCGwJ"),
    "het8_fbc": (lambda: cantilever(8, 2, 2, E=lognormal_E(7, 32, cov=0.3)), False, ["-p", "-fbc"],
                 "femEqsCGwJ"),
    "h40_cg": (lambda: cantilever(40, 8, 8, E=lognormal_E(3, 2560, cov=0.3)), False, ["-p"], "femEqsCGwJ"),
}


@needs_build
@pytest.mark.parametrize("tag", list(CASES))
def test_relinked_fem_program(tag, tmp_path):
    mk, ids, flags, solver = CASES[tag]
    mesh = mk()
    fin = str(tmp_path / "in.fem")
    write_fem_input(mesh, fin, ids=input_ids(mesh, 5) if ids else None)
    out_gpu, out_ref = str(tmp_path / "gpu.res"), str(tmp_path / "ref.res")
    log = run(FEM_GPU, flags + ["-i", fin], out_gpu, env={"UFEM_TRACE": "1"})
    lines = [ln for ln in log.splitlines() if "[ufem]" in ln or "on the GPU" in ln]
    print(tag, *lines, sep="\n  ")
    assert "stiffness matrix on the GPU" in log
    assert f"[ufem] {solver} on the GPU" in log and "on the CPU" not in log
    run(FEM_REF, flags + ["-i", fin], out_ref)
    mine = read_res(out_gpu)
    compare(mine, read_res(out_ref))
    if tag == "het8_fbc":
        compare(mine, read_res(os.path.join(HERE, "golden", "het8_fbc.res")))


@needs_build
def test_relinked_fem_cpu_fill(tmp_path):
    """UFEM_FILL_K=cpu: uFEM's own fem_fill_K assembles K (through the alias
    fem_fill_K_cpu) and the GPU still solves; same result file"""
    mesh = cantilever(20, 4, 4)
    fin = str(tmp_path / "in.fem")
    write_fem_input(mesh, fin)
    log = run(FEM_GPU, ["-p", "-i", fin], str(tmp_path / "a.res"), env={"UFEM_TRACE": "1", "UFEM_FILL_K": "cpu"})
    assert "stiffness matrix on the GPU" not in log and "[ufem] femEqsCGwJ on the GPU" in log
    run(FEM_REF, ["-p", "-i", fin], str(tmp_path / "b.res"))
    compare(read_res(str(tmp_path / "a.res")), read_res(str(tmp_path / "b.res")))
