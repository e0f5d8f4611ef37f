"""The learned preconditioner's kernels against the FP64 oracle
(oracle/unet_oracle.py), and flexible CG with the network.

* single convolution kernels on integer-valued data (exact in bf16, sums
  exact in f32): every mode, the tap flip and the weight gradient must match
  the oracle bit for bit; f32 activations (three-way bf16 split) must match
  the bf16 path on bf16-exact data and the FP64 oracle to 2e-6 otherwise;
* the whole network on bf16-rounded weights: forward within 1e-5 (f32
  arithmetic); parameter gradient within 3e-2 for the weights (the weight
  gradient kernel rounds its operands to bf16) and 1e-4 for the gates;
* the energy loss of stored training pairs within 1e-4 of the oracle's;
* flexible CG: with N = 0 it is Jacobi PCG (iterations +-2); a non-finite
  network output switches it to Jacobi PCG (fallback flag, same solution);
  a trained network needs fewer iterations than Jacobi PCG on a field it
  was not trained on, to the same solution (1e-6 relative)."""

import numpy as np
import pytest
import scipy.sparse as sp

import unet_oracle as U
from ufem_amd import gpu as G
from ufem_amd.mesh import cantilever, grf_E

pytestmark = pytest.mark.gpu


def bf16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def ints(rng, shape, lo=-3, hi=3):
    return rng.integers(lo, hi + 1, size=shape).astype(np.float32)


SHAPES = [((5, 4, 7), 1, 16, 16), ((6, 3, 9), 2, 4, 32), ((3, 5, 5), 1, 48, 16), ((5, 5, 9), 1, 96, 32),
          ((4, 4, 4), 3, 64, 64), ((5, 5, 9), 1, 16, 3), ((5, 5, 9), 2, 32, 64)]


@pytest.mark.parametrize("x_f32", [0, 1, 2])
@pytest.mark.parametrize("shape,nb,cin,cout", SHAPES)
def test_conv_stride1_and_flip_exact(shape, nb, cin, cout, x_f32):
    rng = np.random.default_rng(cin * 7 + cout)
    x = ints(rng, (nb,) + shape + (cin,))
    w = ints(rng, (27, cout, cin), -2, 2) / 4
    for flip in (0, 1):
        y = G.probe_conv(0, flip, x.reshape(nb, -1, cin), w, shape, shape, x_f32=x_f32)
        wr = w[::-1] if flip else w
        ref = np.stack([U.conv(x[b].astype(np.float64), wr.astype(np.float64), 1) for b in range(nb)])
        assert np.array_equal(y.reshape(ref.shape), ref.astype(np.float32)), f"flip={flip}"


@pytest.mark.parametrize("x_f32", [0, 1, 2])
@pytest.mark.parametrize("shape,nb,cin,cout", SHAPES[:4] + SHAPES[6:])
def test_conv_stride2_and_transpose_exact(shape, nb, cin, cout, x_f32):
    rng = np.random.default_rng(cin + 3 * cout)
    cs = U.coarse_shape(shape)
    x = ints(rng, (nb,) + shape + (cin,))
    w = ints(rng, (27, cout, cin), -2, 2) / 4
    y = G.probe_conv(1, 0, x.reshape(nb, -1, cin), w, shape, cs, x_f32=x_f32)
    ref = np.stack([U.conv(x[b].astype(np.float64), w.astype(np.float64), 2) for b in range(nb)])
    assert np.array_equal(y.reshape(ref.shape), ref.astype(np.float32))
    # T2: coarse -> fine, y[f] = sum_t W[t] x[(f - off_t) / 2]: the adjoint of
    # the stride-2 convolution with weights W^T
    xc = ints(rng, (nb,) + cs + (cin,))
    yt = G.probe_conv(2, 0, xc.reshape(nb, -1, cin), w, cs, shape, x_f32=x_f32)
    wT = np.transpose(w, (0, 2, 1)).astype(np.float64)
    reft = np.stack([U.conv_adjoint(xc[b].astype(np.float64), wT, 2, shape) for b in range(nb)])
    assert np.array_equal(yt.reshape(reft.shape), reft.astype(np.float32))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_conv_f32_split_accuracy(mode):
    """f32 activations through the three-way bf16 split: FP64-oracle accuracy
    to 2e-6 relative (bf16-rounded weights are exact in the oracle)."""
    rng = np.random.default_rng(mode)
    fine, cin, cout = (5, 6, 9), 32, 16
    cs = U.coarse_shape(fine)
    gi, go = {0: (fine, fine), 1: (fine, cs), 2: (cs, fine)}[mode]
    x = rng.standard_normal((2,) + gi + (cin,)).astype(np.float32)
    w = bf16(rng.standard_normal((27, cout, cin)) / np.sqrt(27 * cin))
    y = G.probe_conv(mode, 0, x.reshape(2, -1, cin), w, gi, go, x_f32=True)
    w64 = w.astype(np.float64)
    if mode == 2:
        ref = np.stack([U.conv_adjoint(x[b].astype(np.float64), np.transpose(w64, (0, 2, 1)), 2, go)
                        for b in range(2)])
    else:
        ref = np.stack([U.conv(x[b].astype(np.float64), w64, mode + 1) for b in range(2)])
    assert rel(y, ref) <= 2e-6
    # the split stored by the producer gives the same products, bit for bit
    ys = G.probe_conv(mode, 0, x.reshape(2, -1, cin), w, gi, go, x_f32=2)
    assert np.array_equal(ys, y)
    # a bf16-rounded input (the bf16 path) is far less accurate
    yb = G.probe_conv(mode, 0, x.reshape(2, -1, cin), w, gi, go, x_f32=False)
    assert rel(yb, ref) > 1e-4


@pytest.mark.parametrize("stride", [1, 2])
@pytest.mark.parametrize("shape,nb,cin,cout", SHAPES)
def test_weight_gradient_exact(stride, shape, nb, cin, cout):
    rng = np.random.default_rng(5 + cin + cout + stride)
    os_ = shape if stride == 1 else U.coarse_shape(shape)
    x = ints(rng, (nb,) + shape + (cin,), -2, 2)
    w = np.zeros((27, cout, cin), dtype=np.float32)
    dy = ints(rng, (nb,) + os_ + (cout,), -2, 2)
    for x_f32 in (0, 1):
        _, dw = G.probe_conv(stride - 1, 0, x.reshape(nb, -1, cin), w, shape, os_, dy=dy.reshape(nb, -1, cout),
                             x_f32=x_f32)
        ref = sum(U.conv_wgrad(x[b].astype(np.float64), dy[b].astype(np.float64), stride) for b in range(nb))
        assert np.array_equal(dw, ref.astype(np.float32))


def pack(P):
    L = U.n_levels(P)
    return np.concatenate([P[f"W{l}"].ravel() for l in range(L)] +
                          [P[f"{k}{l}"] for l in range(L) for k in ("a", "b")]).astype(np.float32)


def unpack(v, levels=3, fine=True, width=64):
    P, o = {}, 0
    lay = U.layers(levels, fine, width)
    for name, ci, co, _ in lay:
        P[name] = v[o:o + 27 * co * ci].reshape(27, co, ci).astype(np.float64)
        o += 27 * co * ci
    for l in range(levels):
        c = lay[l][2]
        for k in ("a", "b"):
            P[f"{k}{l}"] = v[o:o + c].astype(np.float64)
            o += c
    assert o == v.size
    return P


def random_params(seed, levels=3, fine=True, width=64):
    rng = np.random.default_rng(seed)
    P = U.init_params(seed, levels, fine, width)
    for l in range(levels):
        c = P[f"a{l}"].size
        P[f"a{l}"] = 0.3 * rng.standard_normal(c)
        P[f"b{l}"] = 0.3 * rng.standard_normal(c)
    return unpack(bf16(pack(P)), levels, fine, width)      # weights exactly as the kernels see them


@pytest.fixture(scope="module")
def small_net():
    D, H, W = 5, 5, 9
    return G.GpuNet(W, H, D, batch=2), (D, H, W)


def test_network_forward_matches_oracle(small_net):
    net, shape = small_net
    rng = np.random.default_rng(0)
    P = random_params(4)
    net.set_weights(pack(P))
    x = rng.standard_normal((2,) + shape + (3,))
    e0 = 0.5 * rng.standard_normal((2,) + shape)
    y = net.forward(x.reshape(2, -1, 3), e0.reshape(2, -1))
    for b in range(2):
        ref, _ = U.forward(P, x[b], U.ebar_levels(e0[b].astype(np.float32).astype(np.float64)))
        assert rel(y[b], ref) <= 1e-5


def test_network_gradient_matches_oracle(small_net):
    net, shape = small_net
    rng = np.random.default_rng(1)
    P = random_params(5)
    net.set_weights(pack(P))
    x = rng.standard_normal((2,) + shape + (3,))
    e0 = 0.5 * rng.standard_normal((2,) + shape)
    dy = rng.standard_normal((2,) + shape + (3,))
    g = unpack(net.backward(x.reshape(2, -1, 3), e0.reshape(2, -1), dy.reshape(2, -1, 3)))
    ref = {k: 0.0 for k in P}
    for b in range(2):
        _, c = U.forward(P, x[b], U.ebar_levels(e0[b].astype(np.float32).astype(np.float64)))
        gb = U.backward(P, c, dy[b])
        for k in gb:
            ref[k] = ref[k] + gb[k]
    for k in P:
        a, r = g[k].ravel(), ref[k].ravel()
        cos = float(a @ r / (np.linalg.norm(a) * np.linalg.norm(r)))
        tol = 3e-2 if k.startswith("W") else 1e-4
        assert rel(a, r) <= tol and cos >= 0.999, (k, rel(a, r), cos)


@pytest.mark.parametrize("levels,fine,width", [(2, True, 64), (5, True, 64), (2, False, 64), (4, False, 64),
                                                (4, False, 16), (3, True, 16), (6, False, 32)])
def test_deeper_network_forward_and_gradient_match_oracle(levels, fine, width):
    """the L-level network (down to grids of one node along an axis), the
    coarse-first form and narrower layers against the oracle: forward within
    1e-5, parameter gradient as above"""
    D, H, W = 5, 4, 11
    shape = (D, H, W)
    net = G.GpuNet(W, H, D, batch=2, levels=levels, fine=fine, width=width)
    assert net.P == G.lib().ufem_net_nparams_width(levels, int(fine), width) and net.weights().size == net.P
    if width == 64:
        assert net.P == G.lib().ufem_net_nparams_form(levels, int(fine))
    rng = np.random.default_rng(20 + levels)
    P = random_params(30 + levels, levels, fine, width)
    if width == 64:
        assert U.is_fine(P) == fine
    net.set_weights(pack(P))
    x = rng.standard_normal((2,) + shape + (3,))
    e0 = 0.5 * rng.standard_normal((2,) + shape)
    dy = rng.standard_normal((2,) + shape + (3,))
    y = net.forward(x.reshape(2, -1, 3), e0.reshape(2, -1))
    g = unpack(net.backward(x.reshape(2, -1, 3), e0.reshape(2, -1), dy.reshape(2, -1, 3)), levels, fine, width)
    ref = {k: 0.0 for k in P}
    for b in range(2):
        eb = U.ebar_levels(e0[b].astype(np.float32).astype(np.float64), levels, fine)
        yr, c = U.forward(P, x[b], eb)
        assert rel(y[b], yr) <= 1e-5
        gb = U.backward(P, c, dy[b])
        for k in gb:
            ref[k] = ref[k] + gb[k]
    for k in P:
        a, r = g[k].ravel(), ref[k].ravel()
        cos = float(a @ r / (np.linalg.norm(a) * np.linalg.norm(r)))
        tol = 3e-2 if k.startswith("W") else 1e-4
        assert rel(a, r) <= tol and cos >= 0.999, (k, rel(a, r), cos)
    net.close()


def test_levels_out_of_range():
    with pytest.raises(RuntimeError, match="levels"):
        G.GpuNet(9, 5, 5, levels=9)
    assert G.lib().ufem_net_nparams_levels(1) == -1
    assert G.lib().ufem_net_nparams_levels(3) == G.lib().ufem_net_nparams() == 70640
    with pytest.raises(RuntimeError, match="width"):
        G.GpuNet(9, 5, 5, levels=3, width=6)
    assert G.lib().ufem_net_nparams_width(3, 1, 0) == -1 and G.lib().ufem_net_nparams_width(3, 1, 68) == -1
    # six coarse-first levels of at most 16 channels (the bench's form)
    assert G.lib().ufem_net_nparams_width(6, 0, 16) == 6 * 2 * 16 + 27 * 16 * (3 + 5 * 16)


def beam_mesh(seed, nx=24, ny=6, nz=6):
    return cantilever(nx, ny, nz, L=2.0, B=0.5, H=0.5, E=grf_E(seed, nx, ny, nz, (2.0 / nx, 0.5 / ny, 0.5 / nz)))


@pytest.fixture(scope="module")
def beam():
    m = beam_mesh(3)
    s = G.GpuSystem(m)
    s.assemble(2)
    ref = s.solve("pcg", eps=1e-8)
    return m, s, ref, s.u(), (m.nx + 1, m.ny + 1, m.nz + 1)


def test_energy_loss_matches_oracle(beam):
    m, s, ref, u_ref, (W, H, D) = beam
    net = G.GpuNet(W, H, D)
    net.reserve(8)
    net.bind(s, m.E)
    P = random_params(7)
    net.set_weights(pack(P))
    n_added, _ = net.collect(4, eps=1e-11, every=8)
    assert n_added >= 3
    ip, cols, vals = s.K_csr()
    K = sp.csr_matrix((vals, cols, ip), shape=(s.n, s.n))
    dinv = 1.0 / K.diagonal()
    fixed = np.unique(3 * m.nl_node[m.nl_type == 1] + m.nl_dir[m.nl_type == 1] - 1)
    mask = np.ones(s.n)
    mask[fixed] = 0.0
    eb = U.ebar_levels(U.node_log_modulus(m.E, m.nx, m.ny, m.nz).astype(np.float32).astype(np.float64))
    losses = net.eval(0, n_added)
    for i in range(n_added):
        v, rr, ee, er, f = net.pair(i)
        assert f == 0
        assert abs(er - float(ee.astype(np.float64) @ rr.astype(np.float64))) <= 1e-5 * er
        L, _, _ = U.energy_loss(P, K, rr.astype(np.float64), ee.astype(np.float64), dinv, mask, eb, (D, H, W))
        assert abs(losses[i] - L) <= 1e-4 * max(abs(L), 1e-3), (i, losses[i], L)


def test_zero_network_is_jacobi_pcg(beam):
    m, s, ref, u_ref, (W, H, D) = beam
    net = G.GpuNet(W, H, D)
    net.bind(s, m.E)
    net.set_weights(np.zeros(net.P, dtype=np.float32))
    r = s.solve("learned", eps=1e-8, net=net)
    assert r["rv"] == 0 and not r["fallback"]
    assert abs(r["iters"] - ref["iters"]) <= 2
    assert rel(s.u(), u_ref) <= 1e-6


def test_zero_network_learned_mixed_is_refreshed_jacobi_pcg(beam):
    """q = A p from the FP32 copy of A: the FP64 residual is recomputed each
    time the running one drops by 0.1 and before convergence is accepted, so
    the true residual meets the tolerance and u matches FP64 PCG.  The FP32
    matrix costs iterations on this beam (a CPU restatement of the same
    refreshed CG needs 541 against FP64's 265): K u cancels strongly for the
    smooth bending field, so b - K32 u stalls near 1e-5 and each refresh
    starts a correction of that size."""
    m, s, ref, u_ref, (W, H, D) = beam
    net = G.GpuNet(W, H, D)
    net.bind(s, m.E)
    net.set_weights(np.zeros(net.P, dtype=np.float32))
    r = s.solve("learned_mixed", eps=1e-8, net=net)
    assert r["rv"] == 0 and not r["fallback"]
    assert r["true_relres"] <= 1e-8 * (1 + 1e-9)
    assert r["outer"] >= 7                       # about one refresh per decade down to 1e-8
    assert ref["iters"] <= r["iters"] <= 2.2 * ref["iters"]
    assert rel(s.u(), u_ref) <= 1e-6


@pytest.mark.parametrize("solver", ["learned", "learned_mixed"])
def test_non_finite_network_falls_back_to_jacobi(beam, solver):
    m, s, ref, u_ref, (W, H, D) = beam
    net = G.GpuNet(W, H, D)
    net.bind(s, m.E)
    P = random_params(8)
    P["b0"][:] = 200.0                    # exp(200) overflows f32: the output is not finite
    net.set_weights(pack(P))
    r = s.solve(solver, eps=1e-8, net=net)
    assert r["rv"] == 0 and r["fallback"] == 1
    if solver == "learned":
        assert r["outer"] == 0                       # fell back at iteration 0
        assert abs(r["iters"] - ref["iters"]) <= 2
    else:
        assert r["true_relres"] <= 1e-8 * (1 + 1e-9)
        assert ref["iters"] <= r["iters"] <= 2.2 * ref["iters"]
    assert rel(s.u(), u_ref) <= 1e-6


def test_trained_network_beats_jacobi(beam):
    """Pairs from plain-PCG solves on 8 fresh random fields, 600 Adam steps;
    on a field not seen in training, flexible CG with the network needs at
    most 70 % of Jacobi PCG's iterations and reaches the same solution."""
    m, s, ref, u_ref, (W, H, D) = beam
    net = G.GpuNet(W, H, D, batch=4)
    net.init(1)
    net.reserve(256)
    for f in range(8):
        mf = beam_mesh(100 + f)
        s.set_E(mf.E)
        s.fill_K(2)
        s.add_disps()
        net.bind(s, mf.E)
        net.collect(32, eps=1e-11, every=4)
    assert net.npairs() >= 150
    loss, ms = net.train(600, batch=4, lr=3e-3, seed=0)
    assert np.mean(loss[-50:]) < np.mean(loss[:20])
    s.set_E(m.E)
    s.fill_K(2)
    s.add_disps()
    net.bind(s, m.E)
    plain = s.solve("pcg", eps=1e-8)
    u_plain = s.u()
    r = s.solve("learned", eps=1e-8, net=net)
    print(f"plain {plain['iters']} it, learned {r['iters']} it (fallback {r['fallback']}), "
          f"loss {loss[:20].mean():.3g} -> {loss[-50:].mean():.3g}, train {ms:.0f} ms")
    assert r["rv"] == 0 and not r["fallback"] and r["true_relres"] <= 2e-8
    assert rel(s.u(), u_plain) <= 1e-6
    assert r["iters"] <= 0.7 * plain["iters"]
    rm = s.solve("learned_mixed", eps=1e-8, net=net)
    print(f"learned + mixed {rm['iters']} it, {rm['outer']} FP64 refreshes")
    assert rm["rv"] == 0 and not rm["fallback"] and rm["true_relres"] <= 1e-8 * (1 + 1e-9)
    assert rel(s.u(), u_plain) <= 1e-6
    assert rm["iters"] <= 2.2 * r["iters"]


@pytest.mark.parametrize("width", [64, 16])
@pytest.mark.parametrize("solver", ["learned", "learned_mixed"])
def test_fused_last_layer_matches_the_network_stage(beam, solver, width, monkeypatch):
    """The coarse-first network's last layer y = C0^T t0 folded into the
    flexible-CG combine (f32 on the vector ALUs) against the network's own
    stage followed by the plain combine: x after a few iterations agrees to
    f32 rounding (a wrong tap, weight or parity would move it at O(1)), and
    both reach the same solution.  Both on the matrix-core path
    (net_prec f32)."""
    m, s, ref, u_ref, (W, H, D) = beam
    net = G.GpuNet(W, H, D, levels=4, fine=False, width=width)   # C0 with 32 or 16 channels
    net.init(5)
    net.bind(s, m.E)
    for maxit in (3, 11):
        monkeypatch.setenv("UFEM_NO_FUSED_C0T", "1")
        a = s.solve(solver, eps=1e-30, maxit=maxit, net=net, check=False, net_prec="f32")
        ua = s.u()
        monkeypatch.delenv("UFEM_NO_FUSED_C0T")
        b = s.solve(solver, eps=1e-30, maxit=maxit, net=net, check=False, net_prec="f32")
        ub = s.u()
        assert a["iters"] == b["iters"] == maxit and not a["fallback"] and not b["fallback"]
        assert rel(ub, ua) <= 1e-5, (maxit, rel(ub, ua))
    r = s.solve(solver, eps=1e-8, net=net, net_prec="f32")
    assert r["rv"] == 0 and not r["fallback"] and r["true_relres"] <= 1e-8 * (1 + 1e-9)
    assert rel(s.u(), u_ref) <= 1e-6


# ---- FP64 inference (the solver's path for coarse-first networks of 16 or 32
# channels): the same operator with FP64 activations and sums


def exact_gate_params(seed, levels, width):
    """random bf16 weights scaled by 1/4 and every gate exactly 1 (a = -1,
    b = 0: g = exp(0)), so the device's f32 gates equal the oracle's and the
    FP64 path can be held to FP64 rounding"""
    P = random_params(seed, levels, False, width)
    for l in range(levels):
        P[f"W{l}"] = P[f"W{l}"] * 0.25
        P[f"a{l}"][:] = -1.0
        P[f"b{l}"][:] = 0.0
    return P


def beam_ebar(m, levels):
    e0 = U.node_log_modulus(m.E, m.nx, m.ny, m.nz).astype(np.float32).astype(np.float64)
    return U.ebar_levels(e0, levels, fine=False)


@pytest.mark.parametrize("levels,width", [(4, 16), (6, 16), (3, 32), (5, 32)])
def test_fp64_inference_matches_oracle(beam, levels, width):
    """N(v) by the FP64 kernels against the oracle's FP64 forward: 1e-13 with
    exact gates; with general gates only the f32 gate values differ (1e-6)"""
    m, s, ref, u_ref, (W, H, D) = beam
    net = G.GpuNet(W, H, D, levels=levels, fine=False, width=width)
    net.bind(s, m.E)
    eb = beam_ebar(m, levels)
    v = np.random.default_rng(levels + width).standard_normal((D, H, W, 3))
    P = exact_gate_params(11, levels, width)
    net.set_weights(pack(P))
    y = net.apply64(v)
    y_ref, _ = U.forward(P, v, eb)
    print("exact gates", rel(y, y_ref))
    assert rel(y, y_ref) <= 1e-13
    P2 = random_params(12, levels, False, width)
    net.set_weights(pack(P2))
    y2 = net.apply64(v)
    print("general gates", rel(y2, U.forward(P2, v, eb)[0]))
    assert rel(y2, U.forward(P2, v, eb)[0]) <= 1e-6


def cpu_fcg1(K, b, prec, eps, maxit):
    """flexible CG with one kept direction (beta = -(z, q_old) / (p_old, q_old)),
    the recurrences of the device solver, in FP64 on the CPU"""
    x = np.zeros_like(b)
    r = b.copy()
    bn = np.linalg.norm(b)
    p = q = None
    it = 0
    while it < maxit and np.linalg.norm(r) > eps * bn:
        z = prec(r)
        w = K @ z
        if p is None:
            p, q = z, w
        else:
            beta = -(z @ q) / (p @ q)
            p, q = z + beta * p, w + beta * q
        a = (p @ r) / (p @ q)
        x, r = x + a * p, r - a * q
        it += 1
    return x, it


def test_fp64_learned_solve_follows_cpu_fcg(beam):
    """The learned solve with FP64 inference (fused last layer) and the same
    flexible CG on the CPU with the oracle's preconditioner z = dinv r +
    mask sqrt(dinv) N(mask sqrt(dinv) r): the same iteration count (+-1) and
    solution; after 8 iterations the iterates agree to 1e-10, which a wrong
    tap, channel or scale in any stage would break"""
    m, s, ref, u_ref, (W, H, D) = beam
    levels, width = 6, 16
    net = G.GpuNet(W, H, D, levels=levels, fine=False, width=width)
    net.bind(s, m.E)
    P = exact_gate_params(13, levels, width)
    net.set_weights(pack(P))
    ip, cols, vals = s.K_csr()
    K = sp.csr_matrix((vals, cols, ip), shape=(s.n, s.n))
    b = s.F()
    dinv = 1.0 / K.diagonal()
    fixed = np.unique(3 * m.nl_node[m.nl_type == 1] + m.nl_dir[m.nl_type == 1] - 1)
    mask = np.ones(s.n)
    mask[fixed] = 0.0
    eb = beam_ebar(m, levels)

    def prec(r):
        return U.precondition(P, r, dinv, mask, eb, (D, H, W))[0]

    g8 = s.solve("learned", eps=1e-30, maxit=8, net=net, check=False)
    x8, _ = cpu_fcg1(K, b, prec, 1e-30, 8)
    assert g8["iters"] == 8 and not g8["fallback"]
    print("8 iterations: GPU vs CPU", rel(s.u(), x8))
    assert rel(s.u(), x8) <= 1e-10
    r = s.solve("learned", eps=1e-10, rule=1, net=net)
    xc, itc = cpu_fcg1(K, b, prec, 1e-10, 2000)
    print("learned FP64", r["iters"], "CPU", itc)
    assert r["rv"] == 0 and not r["fallback"]
    assert abs(r["iters"] - itc) <= 1
    assert rel(s.u(), xc) <= 1e-8


def test_fp64_learned_solve_is_scale_invariant(beam):
    """The network is linear: loads scaled by 1.1 change only rounding, and
    with FP64 inference the iteration count stays (+-1; the matrix-core path
    moves by up to a third on the 1M-DOF cantilever).  Not asked of the
    mixed form: when its FP64 refreshes fall depends on rounding."""
    solver = "learned"
    m, s, ref, u_ref, (W, H, D) = beam
    levels, width = 6, 16
    net = G.GpuNet(W, H, D, levels=levels, fine=False, width=width)
    net.init(7)
    its = []
    for scale in (1.0, 1.1, 0.9):
        mk = cantilever(m.nx, m.ny, m.nz, L=2.0, B=0.5, H=0.5, E=m.E, total_load=-10000.0 * scale)
        sk = G.GpuSystem(mk)
        sk.assemble(2)
        net.bind(sk, mk.E)
        r = sk.solve(solver, eps=1e-8, net=net)
        assert r["rv"] == 0 and not r["fallback"]
        its.append(r["iters"])
        net.bind(s, m.E)
        sk.close()
    print(solver, "iterations at load scales 1, 1.1, 0.9:", its)
    assert max(its) - min(its) <= 1


def test_learned_mixed_with_remainders_tracks_fp64(beam):
    """learned + mixed with the FP32 copy's bf16 remainders (48-bit values):
    close to the iterations of the FP64 learned solve (here 285 against 265;
    a CPU restatement on a 24k-DOF beam with a trained network: 191 against
    190) and well below the FP32 copy's (435), and the FP64 solution"""
    m, s, ref, u_ref, (W, H, D) = beam
    net = G.GpuNet(W, H, D, levels=6, fine=False, width=16)
    net.init(7)
    net.bind(s, m.E)
    r64 = s.solve("learned", eps=1e-8, net=net)
    rm = s.solve("learned_mixed", eps=1e-8, net=net, f32_rem=True)
    assert rm["rv"] == 0 and not rm["fallback"] and rm["true_relres"] <= 1e-8 * (1 + 1e-9)
    assert rel(s.u(), u_ref) <= 1e-6
    r32 = s.solve("learned_mixed", eps=1e-8, net=net)
    print("learned", r64["iters"], "learned + mixed, 48-bit K", rm["iters"], "FP32 K", r32["iters"])
    assert rm["iters"] <= 1.1 * r64["iters"] + 5
    assert rm["iters"] <= 0.8 * r32["iters"]


def test_net_prec_option_and_stage_timer(beam):
    """net_prec: FP64 inference by default without the coarse term (the same
    iterates as asking for it), f32 on request (the same solution); the FP64
    stage timer reports every stage"""
    m, s, ref, u_ref, (W, H, D) = beam
    net = G.GpuNet(W, H, D, levels=6, fine=False, width=16)
    net.init(3)
    net.bind(s, m.E)
    a = s.solve("learned", eps=1e-8, net=net)
    ua = s.u()
    b = s.solve("learned", eps=1e-8, net=net, net_prec="f64")
    assert a["iters"] == b["iters"] and np.array_equal(ua, s.u())
    c = s.solve("learned", eps=1e-8, net=net, net_prec="f32")
    assert c["rv"] == 0 and not c["fallback"] and rel(s.u(), u_ref) <= 1e-6
    us = net.time_stages64(reps=3)
    assert len(us) == 12 and all(v > 0.0 for v in us.values())
