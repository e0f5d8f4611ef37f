"""The oracle's element results (strains and stresses at the Gauss points,
the results pass of e009_stiff Mode=AF_YES, fem/fem_e009.c:417-535) against
what the mathematics fixes: a linear displacement field has the same strain
at every point of every (distorted) brick, sigma = D epsilon, rigid motions
carry no strain, and the root bending stress of the cantilever is of the
beam-theory size.  A swapped shear order and a plane-strain D are caught."""

import numpy as np

import ufem_oracle as uo
from ufem_amd.mesh import cantilever, lognormal_E

NU = 0.2


def iso_D(E, nu):
    d1 = E * (1 - nu) / ((1 + nu) * (1 - 2 * nu))
    d2 = E * nu / ((1 + nu) * (1 - 2 * nu))
    D = np.zeros((6, 6))
    D[:3, :3] = d2
    D[np.arange(3), np.arange(3)] = d1
    D[np.arange(3, 6), np.arange(3, 6)] = E / (2 * (1 + nu))
    return D


def engineering_strain(A):
    """u = A x: epsilon x, y, z and engineering shears yz, zx, xy"""
    return np.array([A[0, 0], A[1, 1], A[2, 2], A[1, 2] + A[2, 1], A[2, 0] + A[0, 2], A[0, 1] + A[1, 0]])


def system_with_u(mesh, A, t=(0.0, 0.0, 0.0)):
    s = uo.OracleSystem(mesh)
    s.assemble(3)
    u = (mesh.coords @ A.T + np.asarray(t)).reshape(-1)
    s.set_u(u)
    return s


def test_linear_field_gives_constant_strain_and_hooke_stress():
    mesh = cantilever(4, 3, 3, L=1.0, B=0.6, H=0.6, distort=0.3, seed=9, E=lognormal_E(2, 36, cov=0.3))
    A = np.array([[1.0, 0.3, -0.2], [0.5, -0.7, 0.4], [0.1, 0.9, 0.6]]) * 1e-3
    for npts in (2, 3):
        res = system_with_u(mesh, A).element_results(npts)
        eps = engineering_strain(A)
        assert np.abs(res[:, :, 6:] - eps).max() <= 1e-12 * np.abs(eps).max()
        for e in range(mesh.n_elems):
            sig = iso_D(mesh.E[e], mesh.nu[e]) @ eps
            assert np.abs(res[e, :, :6] - sig).max() <= 1e-10 * np.abs(sig).max()
    # planted mistakes: the shear order yz, zx, xy taken as xy, yz, zx, and a
    # plane-strain D, both differ from what the oracle returns
    swapped = eps[[0, 1, 2, 5, 3, 4]]
    assert np.abs(res[:, :, 6:] - swapped).max() > 1e-6 * np.abs(eps).max()
    Dps = iso_D(mesh.E[0], mesh.nu[0])
    Dps[2, :] = 0.0
    assert np.abs(res[0, :, :6] - Dps @ eps).max() > 1e-3 * np.abs(res[0, :, :6]).max()


def test_rigid_motion_has_no_strain():
    mesh = cantilever(3, 2, 2, L=1.0, B=0.5, H=0.5, distort=0.2, seed=3)
    W = np.array([[0.0, -0.3, 0.2], [0.3, 0.0, -0.1], [-0.2, 0.1, 0.0]]) * 1e-3   # infinitesimal rotation
    res = system_with_u(mesh, W, t=(1e-3, -2e-3, 5e-4)).element_results(3)
    assert np.abs(res).max() <= 1e-6 * 1e-3 * mesh.E.max()


def test_cantilever_root_bending_stress_is_beam_sized():
    mesh = cantilever(20, 4, 4)
    s = uo.OracleSystem(mesh)
    s.assemble(3)
    rv, it, rr = s.solve(0, 1e-12, 10 * s.n, 1)
    assert rv == 0
    res = s.element_results(3)
    # elements next to the support, top and bottom fibres: |sigma_x| ~ M c / I
    M = 10000.0 * 2.0
    I = 0.4 ** 4 / 12
    beam = M * 0.2 / I
    slab = [20 * (j + 4 * k) for j in range(4) for k in range(4)]   # the elements at x = 0 (x fastest)
    root = np.abs(res[slab, :, 0]).max()
    assert 0.5 * beam <= root <= 2.0 * beam, (root, beam)
