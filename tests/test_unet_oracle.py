"""The FP64 oracle of the learned preconditioner (oracle/unet_oracle.py)
against what the mathematics fixes: the adjoint identity of each convolution
stride, symmetry and positive semi-definiteness of the network map, the
pooling and node-averaging rules, and central finite differences of the
energy loss for every parameter block.  Each check is shown to catch a
planted mistake."""

import numpy as np
import pytest
import scipy.sparse as sp

import unet_oracle as U

RNG = np.random.default_rng(11)


def test_conv_shift_and_stride():
    X = RNG.standard_normal((5, 4, 7, 2))
    W = np.zeros((27, 2, 2))
    W[9 * 2 + 3 * 1 + 0] = np.eye(2)             # dz=+1, dy=0, dx=-1
    ref = np.zeros_like(X)
    ref[:-1, :, 1:] = X[1:, :, :-1]
    assert np.array_equal(U.conv(X, W, 1), ref)
    W0 = np.zeros((27, 2, 2))
    W0[13] = np.eye(2)
    assert np.array_equal(U.conv(X, W0, 2), X[::2, ::2, ::2])


@pytest.mark.parametrize("stride", [1, 2])
@pytest.mark.parametrize("shape", [(5, 4, 7), (6, 3, 8)])
def test_conv_adjoint_and_weight_gradient(stride, shape):
    X = RNG.standard_normal(shape + (3,))
    W = RNG.standard_normal((27, 5, 3))
    out = shape if stride == 1 else U.coarse_shape(shape)
    dY = RNG.standard_normal(out + (5,))
    lhs = np.sum(U.conv(X, W, stride) * dY)
    assert abs(lhs - np.sum(X * U.conv_adjoint(dY, W, stride, shape))) <= 1e-12 * abs(lhs)
    assert abs(lhs - np.sum(W * U.conv_wgrad(X, dY, stride))) <= 1e-12 * abs(lhs)
    # planted mistake: the adjoint taken with the taps reversed
    assert abs(lhs - np.sum(X * U.conv_adjoint(dY, W[::-1].copy(), stride, shape))) > 1e-3 * abs(lhs)


def random_params(seed, gate_scale=0.3):
    rng = np.random.default_rng(seed)
    P = U.init_params(seed)
    for l, c in enumerate(U.CH):
        P[f"a{l}"] = gate_scale * rng.standard_normal(c)
        P[f"b{l}"] = gate_scale * rng.standard_normal(c)
    return P


def levels(shape, seed):
    return U.ebar_levels(0.5 * np.random.default_rng(seed).standard_normal(shape))


@pytest.mark.parametrize("shape", [(5, 5, 9), (4, 6, 7)])
def test_network_is_symmetric_positive_semidefinite(shape):
    P, eb = random_params(1), levels(shape, 2)
    v1, v2 = RNG.standard_normal(shape + (3,)), RNG.standard_normal(shape + (3,))
    n1, n2 = U.forward(P, v1, eb)[0], U.forward(P, v2, eb)[0]
    a, b = np.sum(n1 * v2), np.sum(v1 * n2)
    assert abs(a - b) <= 1e-12 * max(abs(a), 1.0)
    for v in (v1, v2, RNG.standard_normal(shape + (3,))):
        assert np.sum(U.forward(P, v, eb)[0] * v) > 0
    # planted mistake: an untied decoder (W1 perturbed in the up path) breaks symmetry
    y1, c1 = U.forward(P, v1, eb)
    Pb = dict(P)
    Pb["W1"] = P["W1"] + 0.01 * np.random.default_rng(9).standard_normal(P["W1"].shape)
    t0 = c1["g"][0] * c1["h0"] + U.conv_adjoint(c1["t1"], Pb["W1"], 2, shape)
    bad1 = U.conv_adjoint(t0, P["W0"], 1, shape)
    y2, c2 = U.forward(P, v2, eb)
    t0b = c2["g"][0] * c2["h0"] + U.conv_adjoint(c2["t1"], Pb["W1"], 2, shape)
    bad2 = U.conv_adjoint(t0b, P["W0"], 1, shape)
    assert abs(np.sum(bad1 * v2) - np.sum(v1 * bad2)) > 1e-6 * abs(np.sum(bad1 * v2))


def test_pooling_and_node_average():
    e = RNG.standard_normal((5, 4, 7))
    p = U.pool(e)
    assert p.shape == (3, 2, 4)
    assert np.isclose(p[0, 0, 0], e[0:2, 0:2, 0:2].mean())
    assert np.isclose(p[1, 1, 1], e[1:4, 1:4, 1:4].mean())
    assert np.isclose(p[2, 1, 3], e[3:5, 1:4, 5:7].mean())
    nx, ny, nz = 3, 2, 2
    E = 30e9 * np.exp(RNG.standard_normal(nx * ny * nz))
    L = U.node_log_modulus(E, nx, ny, nz)
    Ee = np.log(E.reshape(nz, ny, nx) / 30e9)
    assert np.isclose(L[0, 0, 0], Ee[0, 0, 0])
    assert np.isclose(L[1, 1, 1], Ee[:, :, :2].mean())


def laplace3(shape):
    """SPD 3-component test operator on the grid (7-point Laplacian + shift)"""
    n = int(np.prod(shape))
    idx = np.arange(n).reshape(shape)
    rows, cols, vals = [], [], []
    for ax in range(3):
        a = np.moveaxis(idx, ax, 0)
        rows += [a[1:].ravel(), a[:-1].ravel()]
        cols += [a[:-1].ravel(), a[1:].ravel()]
        vals += [-np.ones(a[1:].size)] * 2
    L1 = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    L1 = L1 + sp.diags(-np.asarray(L1.sum(1)).ravel() + 0.05)
    return sp.kron(L1, sp.eye(3)).tocsr()


def test_energy_loss_gradients_match_finite_differences():
    shape = (5, 5, 9)
    n = 3 * int(np.prod(shape))
    K = laplace3(shape)
    dinv = 1.0 / K.diagonal()
    mask = np.ones(n)
    mask[:9] = 0.0
    e = RNG.standard_normal(n) * mask
    r = K @ e
    P, eb = random_params(3), levels(shape, 4)
    L, G, _ = U.energy_loss(P, K, r, e, dinv, mask, eb, shape)
    rng = np.random.default_rng(5)
    h = 1e-6
    for name in ("W0", "W1", "W2", "a0", "b0", "a1", "b1", "a2", "b2"):
        for _ in range(2):
            idx = tuple(rng.integers(s) for s in P[name].shape)
            Pp = {k: v.copy() for k, v in P.items()}
            Pm = {k: v.copy() for k, v in P.items()}
            Pp[name][idx] += h
            Pm[name][idx] -= h
            fd = (U.energy_loss(Pp, K, r, e, dinv, mask, eb, shape)[0] -
                  U.energy_loss(Pm, K, r, e, dinv, mask, eb, shape)[0]) / (2 * h)
            assert abs(fd - G[name][idx]) <= 1e-5 * abs(fd) + 1e-9, (name, idx, fd, G[name][idx])


def test_energy_loss_value_and_planted_mistake():
    """L = 1 - 2 (z,r)/(e,r) + (z,Kz)/(e,r): with N = 0 it is the Jacobi
    step's relative energy error; a gradient taken without the factor 2
    fails the finite-difference bar of the test above."""
    shape = (4, 4, 5)
    n = 3 * int(np.prod(shape))
    K = laplace3(shape)
    dinv = 1.0 / K.diagonal()
    mask = np.ones(n)
    e = RNG.standard_normal(n)
    r = K @ e
    P = U.init_params(0)
    for k in ("W0", "W1", "W2"):
        P[k] *= 0.0
    L, G, z = U.energy_loss(P, K, r, e, dinv, mask, U.ebar_levels(np.zeros(shape)), shape)
    zj = dinv * r
    assert np.isclose(L, (e - zj) @ (r - K @ zj) / (e @ r), rtol=1e-13)
    assert np.allclose(z, zj)
    assert all(np.all(G[k] == 0) for k in ("W0", "W1", "W2"))


def random_params_levels(seed, levels, gate_scale=0.3, fine=True, width=64):
    rng = np.random.default_rng(seed)
    P = U.init_params(seed, levels, fine, width)
    for l in range(levels):
        c = P[f"a{l}"].size
        P[f"a{l}"] = gate_scale * rng.standard_normal(c)
        P[f"b{l}"] = gate_scale * rng.standard_normal(c)
    return P


@pytest.mark.parametrize("levels,fine,width", [(2, True, 64), (5, True, 64), (2, False, 64), (4, False, 64),
                                                (4, False, 16), (3, True, 16)])
def test_deeper_network_is_symmetric_positive_semidefinite(levels, fine, width):
    """more levels (down to grids of one node along an axis), the
    coarse-first form and narrower layers keep N linear, symmetric and PSD;
    an untied decoder at the deepest level breaks it"""
    shape = (5, 4, 11)
    P = random_params_levels(6, levels, fine=fine, width=width)
    assert U.n_levels(P) == levels and all(P[f"W{l}"].shape[1] <= width for l in range(levels))
    if width == 64:
        assert U.is_fine(P) == fine
    eb = U.ebar_levels(0.5 * np.random.default_rng(7).standard_normal(shape), levels, fine)
    assert len(eb) == levels and eb[-1].shape == tuple(U.coarse_shape(eb[-2].shape))
    assert eb[0].shape == (shape if fine else tuple(U.coarse_shape(shape)))
    v1, v2 = RNG.standard_normal(shape + (3,)), RNG.standard_normal(shape + (3,))
    n1, n2 = U.forward(P, v1, eb)[0], U.forward(P, v2, eb)[0]
    a, b = np.sum(n1 * v2), np.sum(v1 * n2)
    assert abs(a - b) <= 1e-12 * max(abs(a), 1.0)
    assert np.sum(n1 * v1) > 0 and np.sum(n2 * v2) > 0
    # linearity
    n12 = U.forward(P, v1 + 2.0 * v2, eb)[0]
    assert np.allclose(n12, n1 + 2.0 * n2, rtol=1e-11, atol=1e-11 * np.abs(n12).max())
    # planted: the deepest transpose with a perturbed weight
    L = levels
    Pb = {k: v.copy() for k, v in P.items()}
    Pb[f"W{L - 1}"] = P[f"W{L - 1}"] + 0.05 * np.random.default_rng(8).standard_normal(P[f"W{L - 1}"].shape)

    def untied(v):
        _, c = U.forward(P, v, eb)
        t = c[f"t{L - 1}"]
        for l in range(L - 2, -1, -1):
            W = Pb[f"W{l + 1}"] if l == L - 2 else P[f"W{l + 1}"]
            t = c["g"][l] * c[f"h{l}"] + U.conv_adjoint(t, W, 2, c[f"h{l}"].shape[:3])
        return U.conv_adjoint(t, P["W0"], 1 if fine else 2, shape)
    u1, u2 = untied(v1), untied(v2)
    assert abs(np.sum(u1 * v2) - np.sum(v1 * u2)) > 1e-6 * abs(np.sum(u1 * v2))


@pytest.mark.parametrize("fine,width", [(True, 64), (False, 64), (False, 16)])
def test_deeper_network_gradients_match_finite_differences(fine, width):
    shape = (5, 4, 9)
    levels = 4
    n = 3 * int(np.prod(shape))
    K = laplace3(shape)
    dinv = 1.0 / K.diagonal()
    mask = np.ones(n)
    mask[:6] = 0.0
    e = RNG.standard_normal(n) * mask
    r = K @ e
    P = random_params_levels(9, levels, fine=fine, width=width)
    eb = U.ebar_levels(0.5 * np.random.default_rng(10).standard_normal(shape), levels, fine)
    L, G, _ = U.energy_loss(P, K, r, e, dinv, mask, eb, shape)
    rng = np.random.default_rng(11)
    h = 1e-4      # the deep gates' gradients are small: a smaller step is lost in rounding
    for name in [f"W{l}" for l in range(levels)] + [f"{p}{l}" for l in range(levels) for p in "ab"]:
        idx = tuple(rng.integers(s) for s in P[name].shape)
        Pp = {k: v.copy() for k, v in P.items()}
        Pm = {k: v.copy() for k, v in P.items()}
        Pp[name][idx] += h
        Pm[name][idx] -= h
        fd = (U.energy_loss(Pp, K, r, e, dinv, mask, eb, shape)[0] -
              U.energy_loss(Pm, K, r, e, dinv, mask, eb, shape)[0]) / (2 * h)
        assert abs(fd - G[name][idx]) <= 1e-5 * abs(fd) + 1e-9, (name, idx, fd, G[name][idx])
