"""The oracle's restatement of femEqsCGwSSOR (fem/fem_eqs.c:999-1188)
against what the mathematics fixes: its preconditioner is symmetric
Gauss-Seidel in natural order, M^-1 = (D+U)^-1 D (D+L)^-1, so the same PCG
written with sparse triangular solves must take the same iterations and
reach the same x; the converged x is the dense solution; and a plausible
slip (the D between the sweeps dropped) is caught by the iteration count."""

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import ufem_oracle as uo
from ufem_amd.mesh import cantilever, lognormal_E


def csr(s):
    ip, c, v = s.K_csr()
    return sp.csr_matrix((v, c, ip), shape=(s.n, s.n))


def sgs_pcg(K, b, eps, maxit, drop_d=False):
    """PCG as femEqsCGwSSOR writes it (beta = ro / roro, relative rule)"""
    D = sp.diags(K.diagonal())
    lower = (sp.tril(K, -1) + D).tocsr()
    upper = (sp.triu(K, 1) + D).tocsr()
    x = np.zeros_like(b)
    r = b.copy()
    nb = np.linalg.norm(b)
    p = None
    roro = 0.0
    for it in range(1, maxit + 1):
        zz = sla.spsolve_triangular(lower, r, lower=True)
        z = sla.spsolve_triangular(upper, zz if drop_d else D @ zz, lower=False)
        ro = r @ z
        p = z.copy() if it == 1 else z + (ro / roro) * p
        q = K @ p
        alpha = ro / (p @ q)
        x += alpha * p
        r -= alpha * q
        if np.linalg.norm(r) <= eps * nb:
            return x, it
        roro = ro
    return x, maxit


def system(nx=10, ny=2, nz=2, seed=4):
    mesh = cantilever(nx, ny, nz, E=lognormal_E(seed, nx * ny * nz, cov=0.3))
    s = uo.OracleSystem(mesh)
    s.assemble(3)
    return mesh, s


def test_ssor_cg_is_natural_order_symmetric_gauss_seidel_pcg():
    mesh, s = system()
    K, F = csr(s), s.F()
    rv, it, rr = s.solve(3, 1e-10, 10 * s.n, 1)
    assert rv == 0 and rr <= 1e-10
    x_ref, it_ref = sgs_pcg(K, F, 1e-10, 10 * s.n)
    assert abs(it - it_ref) <= 1, (it, it_ref)
    assert np.linalg.norm(s.u() - x_ref) <= 1e-8 * np.linalg.norm(x_ref)
    # planted slip: the D between the two sweeps dropped
    _, it_bad = sgs_pcg(K, F, 1e-10, 10 * s.n, drop_d=True)
    assert abs(it_bad - it_ref) > 5, (it_bad, it_ref)


def test_ssor_cg_converges_to_the_dense_solution_in_fewer_iterations_than_jacobi():
    mesh, s = system(12, 3, 3, seed=6)
    K = csr(s).toarray()
    F = s.F()
    fixed = np.unique(3 * mesh.nl_node[mesh.nl_type == 1] + mesh.nl_dir[mesh.nl_type == 1] - 1)
    free = np.setdiff1d(np.arange(s.n), fixed)
    u_d = np.zeros(s.n)
    u_d[free] = np.linalg.solve(K[np.ix_(free, free)], F[free])
    rv, it_ssor, _ = s.solve(3, 1e-12, 10 * s.n, 1)
    assert rv == 0
    assert np.linalg.norm(s.u() - u_d) <= 1e-9 * np.linalg.norm(u_d)
    s.set_u(np.zeros(s.n))
    rv, it_jac, _ = s.solve(0, 1e-12, 10 * s.n, 1)
    assert rv == 0
    assert it_ssor < 0.75 * it_jac, (it_ssor, it_jac)


def test_ssor_cg_reference_rule_and_zero_load():
    mesh, s = system(6, 2, 2, seed=8)
    rv, it, rr = s.solve(3, 1e-21, s.n, 0)      # femSolve's call: eps 1e-21, maxIt = n
    assert rv in (0, -1) and it >= 1
    mesh.nl_val[:] = 0.0
    z = uo.OracleSystem(mesh)
    z.assemble(3)
    rv, it, rr = z.solve(3, 1e-10, z.n, 1)
    assert rv == 0 and it == 0 and not np.any(z.u())
