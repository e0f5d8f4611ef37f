"""The oracle against what the mathematics fixes, independently of the
reference: exact integrals of the trilinear brick on a box, rigid-body null
space, rank, the constant-strain patch test, a dense direct solve and global
equilibrium.  Each check is also shown to catch a planted mistake."""

import numpy as np
import pytest

import ufem_oracle as uo
from ufem_amd.mesh import cantilever, lognormal_E

CORNERS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)


def box_xyz(a, b, c, x0=(0.0, 0.0, 0.0)):
    return (CORNERS + 1) / 2 * np.array([a, b, c]) + np.array(x0)


def exact_box_Ke(a, b, c, E, nu):
    """K_e of an a x b x c box by exact 1-D integrals of the linear factors:
    K_ab,ij = int lam dNa_i dNb_j + mu dNa_j dNb_i + mu delta_ij grad Na . grad Nb."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    h = (a, b, c)

    def I(sa, sb, i, j):
        """int d_i N_a d_j N_b over the box (i, j in 0..2)."""
        out = 1.0
        for d in range(3):
            if d == i and d == j:
                out *= sa[d] * sb[d] / h[d]
            elif d == i:
                out *= sa[d] / 2.0
            elif d == j:
                out *= sb[d] / 2.0
            else:
                out *= h[d] / 3.0 if sa[d] == sb[d] else h[d] / 6.0
        return out

    K = np.zeros((24, 24))
    for A in range(8):
        for B in range(8):
            sa, sb = CORNERS[A], CORNERS[B]
            g = sum(I(sa, sb, d, d) for d in range(3))
            for i in range(3):
                for j in range(3):
                    K[3 * A + i, 3 * B + j] = lam * I(sa, sb, i, j) + mu * I(sa, sb, j, i) + (mu * g if i == j else 0.0)
    return K


def rigid_modes(xyz):
    """Six rigid-body displacement fields of the 8 nodes (translations and
    infinitesimal rotations u = w x x)."""
    R = []
    for d in range(3):
        t = np.zeros((8, 3))
        t[:, d] = 1.0
        R.append(t.reshape(-1))
    for w in np.eye(3):
        R.append(np.cross(w, xyz).reshape(-1))
    return np.array(R).T


@pytest.mark.parametrize("npts", [2, 3])
@pytest.mark.parametrize("dims", [(1.0, 1.0, 1.0), (0.1, 0.1, 0.1), (0.3, 0.05, 0.2)])
def test_box_element_matches_exact_integrals(npts, dims):
    Ke = uo.element_stiffness(box_xyz(*dims), 30e9, 0.2, npts)
    ref = exact_box_Ke(*dims, 30e9, 0.2)
    assert np.abs(Ke - ref).max() <= 1e-13 * np.abs(ref).max()


def test_exact_integrals_catch_a_swapped_shear_term():
    """Planted mistake: gamma_yz built from the wrong gradients (the kind of
    H-row mix-up the restatement must not have) changes K_e far beyond the
    tolerance used above."""
    dims = (0.3, 0.05, 0.2)
    ref = exact_box_Ke(*dims, 30e9, 0.2)
    Ke = uo.element_stiffness(box_xyz(*dims), 30e9, 0.2, 2)
    bad = Ke.copy()
    bad[[1, 2]] = bad[[2, 1]]          # swap the v and w rows of node 0
    bad[:, [1, 2]] = bad[:, [2, 1]]
    assert np.abs(bad - ref).max() > 1e-3 * np.abs(ref).max()


@pytest.mark.parametrize("npts", [2, 3])
def test_rigid_body_null_space_and_rank(npts):
    rng = np.random.default_rng(3)
    for _ in range(5):
        xyz = box_xyz(0.2, 0.1, 0.15) + rng.uniform(-0.02, 0.02, (8, 3))   # distorted brick
        Ke = uo.element_stiffness(xyz, 21e9, 0.3, npts)
        R = rigid_modes(xyz)
        scale = np.abs(Ke).max() * np.abs(R).max()
        assert np.abs(Ke @ R).max() <= 1e-12 * scale
        assert np.allclose(Ke, Ke.T, rtol=0, atol=1e-12 * np.abs(Ke).max())
        ev = np.linalg.eigvalsh(Ke)
        assert np.sum(ev > 1e-9 * ev.max()) == 18           # 24 DOF - 6 rigid modes
        assert ev.min() > -1e-9 * ev.max()                 # positive semi-definite
        # planted mistake: one node's ux column mis-scaled by 1 % (a wrong
        # weight or Jacobian entry) leaves the translation null space
        bad = Ke.copy()
        bad[:, 0] *= 1.01
        assert np.abs(bad @ R).max() > 1e-4 * scale


def test_quadrature_orders_agree_on_parallelepipeds_only():
    box = box_xyz(0.2, 0.1, 0.15)
    K2, K3 = uo.element_stiffness(box, 1.0, 0.25, 2), uo.element_stiffness(box, 1.0, 0.25, 3)
    assert np.abs(K2 - K3).max() <= 1e-13 * np.abs(K3).max()
    skew = box.copy()
    skew[6] += [0.03, 0.02, 0.01]                          # not a parallelepiped
    K2, K3 = uo.element_stiffness(skew, 1.0, 0.25, 2), uo.element_stiffness(skew, 1.0, 0.25, 3)
    assert np.abs(K2 - K3).max() > 1e-8 * np.abs(K3).max()


def test_patch_test_constant_strain():
    """Constant-strain field on a distorted mesh: every interior node is in
    equilibrium (K u_lin = 0 there), for both quadrature orders."""
    mesh = cantilever(4, 3, 3, L=1.0, B=0.6, H=0.6, distort=0.3, seed=9)
    G = np.array([[1e-3, 2e-4, -3e-4], [5e-4, -2e-3, 1e-4], [-4e-4, 3e-4, 7e-4]])
    ulin = (mesh.coords @ G.T).reshape(-1)
    x = mesh.coords
    interior = (x[:, 0] > 1e-9) & (x[:, 0] < 1 - 1e-9) & (x[:, 1] > 1e-9) & (x[:, 1] < 0.6 - 1e-9) & \
               (x[:, 2] > 1e-9) & (x[:, 2] < 0.6 - 1e-9)
    idx = np.repeat(np.nonzero(interior)[0] * 3, 3) + np.tile([0, 1, 2], interior.sum())
    for npts in (2, 3):
        s = uo.OracleSystem(mesh)
        s.fill_K(npts)
        f = s.spmv(ulin)
        assert np.abs(f[idx]).max() <= 1e-9 * np.abs(f).max()
        # planted mistake: a quadratic field is not reproduced
        uq = (mesh.coords[:, [0]] ** 2 * np.array([1e-3, 0, 0])).reshape(-1)
        assert np.abs(s.spmv(uq)[idx]).max() > 1e-6 * np.abs(s.spmv(uq)).max()


def test_cg_matches_dense_direct_solve_and_equilibrium():
    mesh = cantilever(10, 2, 2, E=lognormal_E(4, 40, cov=0.3))
    s = uo.OracleSystem(mesh)
    s.assemble(3)
    ip, c, v = s.K_csr()
    K = np.zeros((s.n, s.n))
    rows = np.repeat(np.arange(s.n), np.diff(ip))
    K[rows, c] = v
    F = s.F()
    # the reference's windowed column zeroing leaves entries in columns of
    # fixed DOFs; u is 0 there, so the free-DOF block decides the solution
    fixed = np.nonzero(np.isin(np.arange(s.n), 3 * mesh.nl_node[mesh.nl_type == 1] + mesh.nl_dir[mesh.nl_type == 1] - 1))[0]
    free = np.setdiff1d(np.arange(s.n), fixed)
    u_d = np.zeros(s.n)
    u_d[free] = np.linalg.solve(K[np.ix_(free, free)], F[free])
    rv, it, rr = s.solve(0, 1e-15, 10 * s.n, 1)
    assert rv == 0
    assert np.linalg.norm(s.u() - u_d) <= 1e-10 * np.linalg.norm(u_d)
    s.reactions(3)
    R, rn, rd = s.R()
    load = mesh.nl_val[mesh.nl_type == 3].sum()
    assert abs(R[rd == 3].sum() + load) <= 1e-9 * abs(load)
    assert abs(R[rd == 1].sum()) <= 1e-9 * abs(load) and abs(R[rd == 2].sum()) <= 1e-9 * abs(load)
    # moment about the y axis at x = 0: the support moment balances the tip load
    xz = mesh.coords[rn]
    My = (R[rd == 1] * xz[rd == 1, 2]).sum() - (R[rd == 3] * xz[rd == 3, 0]).sum()
    Lx = mesh.coords[:, 0].max()
    assert abs(My - load * Lx) <= 1e-6 * abs(load * Lx)


def test_tip_deflection_order_of_magnitude():
    """Coarse physics: Euler-Bernoulli + Timoshenko tip deflection bounds the
    trilinear-brick answer within the locking range (+-30 %)."""
    mesh = cantilever(20, 4, 4)
    s = uo.OracleSystem(mesh)
    s.assemble(2)
    s.solve(0, 1e-12, 10 * s.n, 1)
    u = s.u().reshape(-1, 3)
    tip = np.abs(mesh.coords[:, 0] - 2.0) < 1e-12
    w = -u[tip, 2].mean()
    P, L, E, nu, A = 1e4, 2.0, 30e9, 0.2, 0.16
    I = 0.4 ** 4 / 12
    G = E / (2 * (1 + nu))
    w_beam = P * L ** 3 / (3 * E * I) + P * L / (5 / 6 * G * A)
    assert 0.7 * w_beam <= w <= 1.05 * w_beam
