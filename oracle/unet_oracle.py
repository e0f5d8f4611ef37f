"""FP64 numpy restatement of the learned preconditioner's network and of its
training loss -- the checker for the hand-written HIP kernels
(ufem_amd/csrc/ugk_unet.hip, ufem_amd/csrc/ufem_net.c).

TEST INFRASTRUCTURE ONLY (imported by tests/ as the oracle).  The network has
no counterpart in the reference; it preconditions the reference's Jacobi-PCG
solve (femEqsCGwJ, fem/fem_eqs.c:88-237).  Definition:

* fields are channels-last on the node grid, [D][H][W][C] = [z][y][x][c];
  level l+1 keeps every second node: D_{l+1} = (D_l + 1) // 2;
* conv(X, W, s): 3x3x3 taps t = 9 (dz+1) + 3 (dy+1) + (dx+1), zero padding,
  no bias, W[t][co][ci]; stride 1: Y[v] = sum_t W[t] X[v + off_t]; stride 2:
  Y[c] = sum_t W[t] X[2c + off_t]; convT is its adjoint;
* symmetric U-Net N of L levels (2..8, 3 by default) with weights W0 (3->16,
  stride 1), W1 (16->32, stride 2), W2 (32->64, stride 2), W3.. (64->64,
  stride 2) and per-channel gate parameters (a_l, b_l):
      h0 = conv(v, W0, 1), h_l = conv(h_{l-1}, W_l, 2)
      g_l = exp(b_l - (1 + a_l) ebar_l)                (per node, per channel)
      t_{L-1} = g_{L-1} h_{L-1}
      t_l = g_l h_l + convT(t_{l+1}, W_{l+1}, 2)
      N(v) = convT(t0, W0, 1)
  so N = sum_l H_l^T diag(g_l) H_l with H_l = conv chains: linear in v,
  symmetric and positive semi-definite; the number of levels is the number
  of W entries of the parameter dict;
* the coarse-first form has no node-grid level: W0 is a stride-2
  convolution 3->32 and the levels are W1 32->64, W2.. 64->64, so level l
  lives on the grid of the node-grid form's level l + 1 and ebar starts
  pooled once; the form is told by W0's output channels (16: node grid,
  32: coarse first);
* ebar_0 = log(E / 30 GPa) averaged over a node's bricks; ebar_{l+1} = mean
  of ebar_l over the 3x3x3 fine nodes around 2c (those inside the grid);
* preconditioner: z = dinv r + mask dsq N(mask dsq r), dsq = sqrt(dinv);
* training loss of a pair (r, e = K^-1 r): energy-norm error of z relative to
  that of e, L = (e - z)^T (r - K z) / (e^T r).
"""

from __future__ import annotations

import numpy as np

CH_ALL = (16, 32, 64, 64, 64, 64, 64, 64)
CI_ALL = (3, 16, 32, 64, 64, 64, 64, 64)


def layers(levels=3, fine=True, width=64):
    """(name, in channels, out channels, stride) of C0 .. C_{L-1}; every
    channel count capped at `width`"""
    if fine:
        return [(f"W{l}", min(CI_ALL[l], width), min(CH_ALL[l], width), 1 if l == 0 else 2) for l in range(levels)]
    ch = tuple(min(c, width) for c in CH_ALL[1:levels + 1])
    ci = (3,) + tuple(ch[:-1])
    return [(f"W{l}", ci[l], ch[l], 2) for l in range(levels)]


def is_fine(P):
    """the form of a full-width parameter set (W0 with 16 outputs: node grid);
    a narrower set says its form through the grids of ebar (forward)"""
    return P["W0"].shape[1] == CH_ALL[0]


def stride0(v, eb):
    """C0's stride: 1 when ebar_0 lies on the node grid of v (node-grid
    form), 2 when it is pooled once (coarse first)"""
    return 1 if tuple(eb[0].shape) == tuple(v.shape[:3]) else 2


CH = CH_ALL[:3]
LAYERS = layers(3)
OFFS = np.array([(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)])
E_REF = 30e9


def coarse_shape(shape):
    return tuple((s + 1) // 2 for s in shape)


def _take(X, iz, iy, ix):
    D, H, Wd = X.shape[:3]
    ok = (iz >= 0) & (iz < D) & (iy >= 0) & (iy < H) & (ix >= 0) & (ix < Wd)
    out = X[np.clip(iz, 0, D - 1), np.clip(iy, 0, H - 1), np.clip(ix, 0, Wd - 1)]
    return out * ok[..., None]


def _grid(shape):
    return np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")


def conv(X, W, stride):
    out_shape = X.shape[:3] if stride == 1 else coarse_shape(X.shape[:3])
    gz, gy, gx = _grid(out_shape)
    Y = np.zeros(out_shape + (W.shape[1],))
    for t, (dz, dy, dx) in enumerate(OFFS):
        Y += _take(X, stride * gz + dz, stride * gy + dy, stride * gx + dx) @ W[t].T
    return Y


def conv_adjoint(Y, W, stride, in_shape):
    """X-gradient of conv: the adjoint map, from the output grid back to in_shape."""
    gz, gy, gx = _grid(Y.shape[:3])
    D, H, Wd = in_shape
    Xp = np.zeros((D + 2, H + 2, Wd + 2, W.shape[2]))
    for t, (dz, dy, dx) in enumerate(OFFS):
        iz, iy, ix = stride * gz + dz, stride * gy + dy, stride * gx + dx
        g = Y @ W[t]
        ok = (iz >= -1) & (iz <= D) & (iy >= -1) & (iy <= H) & (ix >= -1) & (ix <= Wd)
        np.add.at(Xp, (iz[ok] + 1, iy[ok] + 1, ix[ok] + 1), g[ok])
    return Xp[1:-1, 1:-1, 1:-1]


def conv_wgrad(X, dY, stride):
    """d <dY, conv(X, W)> / dW."""
    gz, gy, gx = _grid(dY.shape[:3])
    out = np.zeros((27, dY.shape[3], X.shape[3]))
    for t, (dz, dy, dx) in enumerate(OFFS):
        Xs = _take(X, stride * gz + dz, stride * gy + dy, stride * gx + dx)
        out[t] = dY.reshape(-1, dY.shape[3]).T @ Xs.reshape(-1, X.shape[3])
    return out


def pool(e):
    """mean over the 3x3x3 fine nodes around 2c that lie inside the grid"""
    cs = coarse_shape(e.shape)
    gz, gy, gx = _grid(cs)
    s = np.zeros(cs)
    n = np.zeros(cs)
    for dz, dy, dx in OFFS:
        iz, iy, ix = 2 * gz + dz, 2 * gy + dy, 2 * gx + dx
        ok = (iz >= 0) & (iz < e.shape[0]) & (iy >= 0) & (iy < e.shape[1]) & (ix >= 0) & (ix < e.shape[2])
        s += np.where(ok, e[np.clip(iz, 0, e.shape[0] - 1), np.clip(iy, 0, e.shape[1] - 1),
                            np.clip(ix, 0, e.shape[2] - 1)], 0.0)
        n += ok
    return s / n


def node_log_modulus(E_elem, nx, ny, nz):
    """log(E / 30 GPa) at the nodes: mean over the (up to 8) adjacent bricks."""
    Ee = np.log(np.asarray(E_elem).reshape(nz, ny, nx) / E_REF)
    acc = np.zeros((nz + 1, ny + 1, nx + 1))
    cnt = np.zeros_like(acc)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                acc[dz:dz + nz, dy:dy + ny, dx:dx + nx] += Ee
                cnt[dz:dz + nz, dy:dy + ny, dx:dx + nx] += 1
    return acc / cnt


def ebar_levels(e0, levels=3, fine=True):
    out = [e0 if fine else pool(e0)]
    for _ in range(1, levels):
        out.append(pool(out[-1]))
    return out


def n_levels(P):
    return sum(1 for k in P if k.startswith("W"))


def init_params(seed=0, levels=3, fine=True, width=64):
    rng = np.random.default_rng(seed)
    P = {}
    for name, ci, co, _ in layers(levels, fine, width):
        P[name] = rng.standard_normal((27, co, ci)) / np.sqrt(27 * ci)
    for l, (_, _, co, _) in enumerate(layers(levels, fine, width)):
        P[f"a{l}"] = np.zeros(co)
        P[f"b{l}"] = np.zeros(co)
    return P


def gates(P, eb):
    return [np.exp(P[f"b{l}"] - (1.0 + P[f"a{l}"]) * eb[l][..., None]) for l in range(n_levels(P))]


def forward(P, v, eb):
    """v [D][H][W][3] -> N(v) [D][H][W][3], with the intermediate fields."""
    L = n_levels(P)
    s0 = stride0(v, eb)
    g = gates(P, eb)
    c = {"v": v, "g": g, "eb": eb, "s0": s0}
    h = [conv(v, P["W0"], s0)]
    for l in range(1, L):
        h.append(conv(h[-1], P[f"W{l}"], 2))
    t = [None] * L
    t[L - 1] = g[L - 1] * h[L - 1]
    for l in range(L - 2, -1, -1):
        t[l] = g[l] * h[l] + conv_adjoint(t[l + 1], P[f"W{l + 1}"], 2, h[l].shape[:3])
    for l in range(L):
        c[f"h{l}"], c[f"t{l}"] = h[l], t[l]
    y = conv_adjoint(t[0], P["W0"], s0, v.shape[:3])
    return y, c


def backward(P, c, dy):
    """gradient of <dy, N(v)> with respect to every parameter"""
    L = n_levels(P)
    s0 = c["s0"]
    g = c["g"]
    G = {}
    dt = [conv(dy, P["W0"], s0)]                 # adjoint of y = conv_adjoint(t0)
    G["W0"] = conv_wgrad(dy, c["t0"], s0)
    for l in range(1, L):
        dt.append(conv(dt[-1], P[f"W{l}"], 2))
        G[f"W{l}"] = conv_wgrad(dt[l - 1], c[f"t{l}"], 2)
    dh = g[L - 1] * dt[L - 1]
    for l in range(L - 1, 0, -1):
        G[f"W{l}"] += conv_wgrad(c[f"h{l - 1}"], dh, 2)
        dh = g[l - 1] * dt[l - 1] + conv_adjoint(dh, P[f"W{l}"], 2, c[f"h{l - 1}"].shape[:3])
    G["W0"] += conv_wgrad(c["v"], dh, s0)
    for l in range(L):
        q = g[l] * c[f"h{l}"] * dt[l]
        G[f"b{l}"] = q.reshape(-1, q.shape[-1]).sum(0)
        G[f"a{l}"] = -(q * c["eb"][l][..., None]).reshape(-1, q.shape[-1]).sum(0)
    return G


def precondition(P, r, dinv, mask, eb, shape):
    """z = dinv r + mask dsq N(mask dsq r) for a DOF vector r (DOF 3 node + c)."""
    dsq = np.sqrt(dinv)
    v = (mask * dsq * r).reshape(shape + (3,))
    y, c = forward(P, v, eb)
    return dinv * r + mask * dsq * y.reshape(-1), c


def energy_loss(P, K, r, e, dinv, mask, eb, shape):
    """L = (e - z)^T (r - K z) / (e^T r) and its parameter gradient."""
    z, c = precondition(P, r, dinv, mask, eb, shape)
    res = r - K @ z
    er = float(e @ r)
    L = float((e - z) @ res) / er
    dz = -2.0 * res / er
    dy = (mask * np.sqrt(dinv) * dz).reshape(shape + (3,))
    return L, backward(P, c, dy), z
