"""The two-level correction (ufem_sys_set_coarse, UFEM_SOLVER_TWOLEVEL): the
device's dense SPD inverse against numpy, the coarse matrix it assembles from
the elements' Lame constants against the Galerkin product P^T K P formed here
from the assembled K, and the two-level PCG against the plain solve.

Bars: inverse within 1e-10 (relative to the inverse's size); coarse matrix
within 1e-10 of max|A| (the sums differ only in order); displacements within
1e-8 relative; reactions within 1e-6."""

import numpy as np
import pytest
import scipy.sparse as sp

from ufem_amd import gpu as G
from ufem_amd.mesh import cantilever, lognormal_E

pytestmark = pytest.mark.gpu


def prolongation(W, H, D, cx, cy, cz):
    """trilinear P from the coarse node grid: fine node (i, j, k) in cell
    min(i // cx, Wc - 2), ..., weight t / c for the corner at offset 1"""
    Wc, Hc, Dc = (W - 1) // cx + 1, (H - 1) // cy + 1, (D - 1) // cz + 1
    rows, cols, vals = [], [], []
    for k in range(D):
        ck = min(k // cz, Dc - 2)
        tz = (k - ck * cz) / cz
        for j in range(H):
            cj = min(j // cy, Hc - 2)
            ty = (j - cj * cy) / cy
            for i in range(W):
                ci = min(i // cx, Wc - 2)
                tx = (i - ci * cx) / cx
                node = i + W * (j + H * k)
                for dz in (0, 1):
                    for dy in (0, 1):
                        for dx in (0, 1):
                            w = (tx if dx else 1 - tx) * (ty if dy else 1 - ty) * (tz if dz else 1 - tz)
                            if w == 0.0:
                                continue
                            I = (ci + dx) + Wc * ((cj + dy) + Hc * (ck + dz))
                            for a in range(3):
                                rows.append(3 * node + a)
                                cols.append(3 * I + a)
                                vals.append(w)
    return sp.csr_matrix((vals, (rows, cols)), shape=(3 * W * H * D, 3 * Wc * Hc * Dc)), (Wc, Hc, Dc)


@pytest.mark.parametrize("n", [64, 100, 200])
def test_dense_spd_inverse(n):
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n))
    A = B @ B.T + n * np.eye(n) * 0.1
    A[: n // 3] *= 1e3                       # a spread of scales, kept symmetric
    A[:, : n // 3] *= 1e3
    inv = G.dense_spd_inverse(A)
    ref = np.linalg.inv(A)
    assert np.abs(inv - ref).max() <= 1e-10 * np.abs(ref).max()
    assert np.abs(A @ inv - np.eye(n)).max() <= 1e-9


def test_coarse_matrix_is_the_galerkin_product():
    nx, ny, nz, cx, cy, cz = 16, 4, 4, 4, 2, 2
    mesh = cantilever(nx, ny, nz, E=lognormal_E(5, nx * ny * nz, cov=0.4))
    rng = np.random.default_rng(1)
    mesh.nu[:] = rng.uniform(0.15, 0.3, mesh.n_elems)   # lambda and mu both vary
    g = G.GpuSystem(mesh)
    g.assemble(2)
    nc = g.set_coarse(nx + 1, ny + 1, nz + 1, cx, cy, cz)
    ip, cols, vals = g.K_csr()
    K = sp.csr_matrix((vals, cols, ip), shape=(g.n, g.n))
    P, (Wc, Hc, Dc) = prolongation(nx + 1, ny + 1, nz + 1, cx, cy, cz)
    assert nc == P.shape[1]
    fixed = 3 * mesh.nl_node[mesh.nl_type == 1] + mesh.nl_dir[mesh.nl_type == 1] - 1
    free = np.ones(g.n, bool)
    free[fixed] = False
    coinc = np.array([3 * ((I % Wc) * cx + (nx + 1) * (((I // Wc) % Hc) * cy + (ny + 1) * (I // (Wc * Hc)) * cz)) + a
                      for I in range(Wc * Hc * Dc) for a in range(3)])
    act = free[coinc]
    Pa = P[:, act]
    ref = np.eye(nc)
    ref[np.ix_(act, act)] = (Pa.T @ K @ Pa).toarray()
    A = g.coarse_matrix()
    assert np.abs(A - ref).max() <= 1e-10 * np.abs(ref).max()
    Ainv = g.coarse_matrix(inverse=True)
    assert np.abs(Ainv @ ref - np.eye(nc)).max() <= 1e-8
    # the weights matter: P with each cell's weights mirrored gives another matrix
    Pm, _ = prolongation(nx + 1, ny + 1, nz + 1, cx, cy, cz)
    Pm = Pm[:, act]
    Pm.data = 1.0 - Pm.data
    assert np.abs((Pm.T @ K @ Pm).toarray() - ref[np.ix_(act, act)]).max() > 1e-3 * np.abs(ref).max()
    g.close()


def test_twolevel_solve_matches_plain_and_cuts_iterations():
    nx, ny, nz = 48, 8, 8
    mesh = cantilever(nx, ny, nz, E=lognormal_E(11, nx * ny * nz, cov=0.3))
    g = G.GpuSystem(mesh)
    g.assemble(2)
    ref = g.solve("pcg", eps=1e-12, rule=1)
    assert ref["rv"] == 0
    u0 = g.u()
    R0 = g.reactions()[0]
    g.set_coarse(nx + 1, ny + 1, nz + 1, 8, 8, 8)
    info = g.solve("twolevel", eps=1e-12, rule=1)
    print("jacobi", ref["iters"], "two-level", info["iters"], info["solve_ms"])
    assert info["rv"] == 0, G.last_error()
    u = g.u()
    assert np.linalg.norm(u - u0) <= 1e-8 * np.linalg.norm(u0)
    assert np.linalg.norm(g.reactions()[0] - R0) <= 1e-6 * np.linalg.norm(R0)
    assert info["true_relres"] <= 1e-10
    assert info["iters"] * 3 < ref["iters"]
    # the same result again (the coarse matrix rebuilt, same sums)
    again = g.solve("twolevel", eps=1e-12, rule=1)
    assert again["iters"] == info["iters"] and np.array_equal(g.u(), u)
    # q = K p from the FP32 copy with FP64 refreshes: the true residual meets
    # the tolerance (accepted on a refreshed residual only)
    mx = g.solve("twolevel_mixed", eps=1e-10, rule=1)
    print("two-level mixed", mx["iters"], "refreshes", mx["outer"], mx["true_relres"])
    assert mx["rv"] == 0, G.last_error()
    assert mx["true_relres"] <= 1e-10 and mx["outer"] >= 1
    assert np.linalg.norm(g.u() - u0) <= 1e-7 * np.linalg.norm(u0)
    assert mx["iters"] <= 1.5 * info["iters"]
    g.close()


def test_coarse_in_flexible_cg():
    """the correction inside flexible CG: without a network (FP64 and with the
    FP32 matrix), and added to a network's term; all reach the plain solve's u"""
    nx, ny, nz = 32, 8, 8
    mesh = cantilever(nx, ny, nz, E=lognormal_E(13, nx * ny * nz, cov=0.3))
    g = G.GpuSystem(mesh)
    g.assemble(2)
    ref = g.solve("pcg", eps=1e-12, rule=1)
    u0 = g.u()
    g.set_coarse(nx + 1, ny + 1, nz + 1, 8, 4, 4)
    tl = g.solve("twolevel", eps=1e-12, rule=1)
    net = G.GpuNet(nx + 1, ny + 1, nz + 1, batch=1)
    net.init(3)
    net.bind(g, mesh.E)
    # the FP32-matrix form is held to 1e-10 (its refreshes stall short of 1e-12)
    for kw, eps, tol in ((dict(solver="learned", coarse=True), 1e-12, 1e-8),
                         (dict(solver="learned_mixed", coarse=True), 1e-10, 1e-7),
                         (dict(solver="learned", coarse=True, net=net, fcg_m=1), 1e-12, 1e-8)):
        r = g.solve(eps=eps, rule=1, **kw)
        print(kw.get("solver"), "net" if "net" in kw else "", r["iters"], tl["iters"], ref["iters"])
        assert r["rv"] == 0, (kw, G.last_error())
        assert np.linalg.norm(g.u() - u0) <= tol * np.linalg.norm(u0)
        assert r["iters"] < ref["iters"] / 2
    # the coarse flag needs a coarse grid, and the network-free form needs it
    h = G.GpuSystem(mesh)
    h.assemble(2)
    assert h.solve("learned", eps=1e-8, rule=1, coarse=True, check=False)["rv"] == G.AF_ERR_VAL
    assert h.solve("learned", eps=1e-8, rule=1, check=False)["rv"] == G.AF_ERR_VAL
    net.close()
    h.close()
    g.close()


def test_twolevel_errors():
    mesh = cantilever(8, 4, 4, distort=0.2, seed=3)
    g = G.GpuSystem(mesh)
    g.assemble(2)
    with pytest.raises(RuntimeError):
        g.set_coarse(9, 5, 5, 4, 2, 2)       # not a regular grid
    with pytest.raises(RuntimeError):
        g.set_coarse(9, 5, 5, 3, 2, 2)       # 3 does not divide 8
    r = g.solve("twolevel", eps=1e-8, rule=1, check=False)
    assert r["rv"] == G.AF_ERR_VAL           # no coarse grid
    g.close()
