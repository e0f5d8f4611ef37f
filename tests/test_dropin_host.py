"""Host-side behaviour of the drop-in boundary that needs no GPU: calls with
types the GPU does not serve go to the reference's CPU functions under their
renamed symbols (<function>_cpu) when the program provides them and return
AF_ERR_TYP otherwise; a zero load returns x = 0 before anything else, as
femEqsCGwJ does (fem/fem_eqs.c:103-116); femMatSetZeroBig/femVecSetZeroBig
clear every stored value (fem/fem_math.c:483-523, :920-958).

The fallback is checked in child processes: a stub library standing in for
the renamed reference objects is loaded (RTLD_GLOBAL) before libufem_gpu.so."""

import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ufem_amd", "libufem_gpu.so")

STUB = r"""
typedef struct { long type, rows, cols, len; long *pos; double *data; long *frompos, *defpos; } tMatrix;
typedef struct { long type, rows, len; long *pos; double *data; } tVector;
int femEqsCGwJ_cpu(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt)
{ x->data[0] = 42.0 + eps + (double)maxIt + (double)a->type; (void)b; return 77; }
int femEqsCGwSSOR_cpu(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt)
{ (void)a; (void)b; (void)eps; (void)maxIt; x->data[0] = 5.0; return 78; }
double femVecNormBig_cpu(tVector *a) { return 3.0 * (double)a->type; }
int femMatVecMultBig_cpu(tMatrix *a, tVector *b, tVector *c)
{ (void)a; (void)b; c->data[0] = -1.0; return 79; }
"""

CHILD = textwrap.dedent(r"""
    import ctypes as C, sys
    sys.path.insert(0, {root!r})
    if {stub!r}:
        C.CDLL({stub!r}, mode=C.RTLD_GLOBAL)
    from ufem_amd import gpu as G
    import numpy as np
    L = G.lib()
    n = 3
    data = np.array([4.0, 1, 0, 1, 4, 1, 0, 1, 4])
    A = G.tMatrix(0, n, n, 9, None, G._pd(data), None, None)        # MAT_FULL
    b = G.Vec(np.ones(n)); x = G.Vec(np.zeros(n))
    print(L.femEqsCGwJ(C.byref(A), C.byref(b.s), C.byref(x.s), 0.5, 10), x.data[0])
    print(L.femEqsCGwSSOR(C.byref(A), C.byref(b.s), C.byref(x.s), 0.5, 10), x.data[0])
    y = G.Vec(np.zeros(n))
    print(L.femMatVecMultBig(C.byref(A), C.byref(b.s), C.byref(y.s)), y.data[0])
    sv = G.tVector(1, n, n, None, G._pd(np.ones(n)))                 # VEC_SPAR
    print(L.femVecNormBig(C.byref(sv)))
""")


def _run(stub):
    code = CHILD.format(root=ROOT, stub=stub)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, HIP_VISIBLE_DEVICES=""))
    assert out.returncode == 0, out.stderr
    return [ln.split() for ln in out.stdout.strip().splitlines()]


@pytest.fixture(scope="module")
def stub_lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("stub")
    src, so = d / "stub.c", d / "libfem_cpu_stub.so"
    src.write_text(STUB)
    subprocess.run(["gcc", "-shared", "-fPIC", "-O1", "-o", str(so), str(src)], check=True)
    return str(so)


def test_unserved_types_go_to_the_cpu_functions(stub_lib):
    r = _run(stub_lib)
    assert r[0] == ["77", str(42.0 + 0.5 + 10 + 0)]
    assert r[1] == ["78", "5.0"]
    assert r[2] == ["79", "-1.0"]
    assert float(r[3][0]) == 3.0


def test_unserved_types_without_cpu_functions():
    r = _run("")
    assert r[0][0] == "-5" and r[1][0] == "-5" and r[2][0] == "-5"
    assert r[3][0] == "nan"


def test_zero_load_and_set_zero_need_no_gpu():
    import ctypes as C
    from ufem_amd import gpu as G
    n = 6
    rl = np.full(n, 1)
    K = G.RowSlotMatrix(n, rl, np.arange(1, n + 1), np.full(n, 2.0), capacity=rl + 1)
    b = G.Vec(np.zeros(n))
    x = G.Vec(np.full(n, 7.0))
    assert G.femEqsCGwJ(K, b, x, 1e-8, 10) == 0
    assert not x.data.any()
    # a size mismatch is reported before the zero-load exit
    assert G.lib().femEqsCGwJ(C.byref(K.s), C.byref(G.Vec(np.zeros(n - 1)).s), C.byref(x.s), 1e-8, 10) == -9
    K.set_zero()
    assert not K.data.any() and K.pos[:n].tolist() == [1, 0, 2, 0, 3, 0]
    v = G.Vec(np.arange(5.0))
    G.lib().femVecSetZeroBig(C.byref(v.s))
    assert not v.data.any()
