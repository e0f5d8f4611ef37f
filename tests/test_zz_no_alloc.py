"""ufem_sys_solve_noalloc (the solve entry point that allocates nothing; its
workspace is a fixed device array) against ufem_sys_solve on the same
systems: two calls in a row, the second on a smaller system, each equal to
the allocating entry point's result bit for bit and to the oracle within
1e-8; sizes beyond the workspace and other solvers are refused."""

import ctypes as C

import numpy as np
import pytest

import ufem_oracle as uo
from ufem_amd import gpu as G
from ufem_amd.mesh import cantilever, lognormal_E

pytestmark = pytest.mark.gpu


def _solve(g, noalloc, solver="pcg", eps=1e-12):
    o = G.SolveOpts(G.SOLVER[solver], 1, eps, 0, G.EXEC["streamed"], 0, 0.0, 0, None, 0, 0, 0, 0)
    info = G.SolveInfo()
    fn = G.lib().ufem_sys_solve_noalloc if noalloc else G.lib().ufem_sys_solve
    rv = fn(g.h, C.byref(o), C.byref(info))
    return rv, info.as_dict(), g.u()


def test_noalloc_equals_allocating_solve_twice():
    meshes = [cantilever(40, 8, 8, E=lognormal_E(31, 2560, cov=0.3)),   # 9,963 unknowns
              cantilever(12, 3, 3, E=lognormal_E(32, 108, cov=0.3))]    # then 624
    for mesh in meshes:
        g = G.GpuSystem(mesh)
        g.assemble(2)
        rv0, i0, u0 = _solve(g, False)
        rv1, i1, u1 = _solve(g, True)
        print(f"n {g.n}: allocating {i0['iters']} it, no-alloc {i1['iters']} it, true {i1['true_relres']:.2e}")
        assert rv0 == 0 and rv1 == 0
        assert i1["iters"] == i0["iters"]
        assert np.array_equal(u1, u0)
        assert abs(i1["true_relres"] - i0["true_relres"]) <= 1e-12 * max(i0["true_relres"], 1e-300) + 1e-30
        o = uo.OracleSystem(mesh)
        o.assemble(2)
        o.solve(0, 1e-15, 20 * o.n, 1)
        assert np.linalg.norm(u1 - o.u()) <= 1e-8 * np.linalg.norm(o.u())
        g.close()


def test_noalloc_refuses_what_it_cannot_hold():
    g = G.GpuSystem(cantilever(8, 2, 2))
    g.assemble(2)
    rv, _, _ = _solve(g, True, solver="mixed")
    assert rv == G.AF_ERR_TYP and "streamed form only" in G.last_error()
    g.close()
    big = cantilever(216, 40, 40)          # 1,094,331 unknowns > 2^20
    g = G.GpuSystem(big)
    g.assemble(2)
    rv, _, _ = _solve(g, True)
    assert rv == -11 and "at most 1048576" in G.last_error()
    g.close()
