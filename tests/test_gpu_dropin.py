"""The drop-in boundary on the GPU: the device copies the drop-in keeps of
uFEM's row-slot matrices (3x3 node blocks or scalar rows, refresh and
validity policy), femMatSetZeroBig, and fem_fill_K(AF_NO) on the GPU writing
into the reference's row-slot K through its DOF map, with and without -fbc.

References: golden vectors written by the compiled reference
(tests/golden/ufem_golden.npz: c1_cg, het8_cg, het8_fbc, pd_fbc) and the
oracle restatement.  Bars: K within 1e-12 of max|K| with the identical
pattern; u within 1e-8 relative."""

import numpy as np
import pytest

import ufem_oracle as uo
from tests.golden.make_golden import pattern_digest, probe_vector, csr_matvec
from ufem_amd import gpu as G
from ufem_amd.mesh import cantilever, lognormal_E, prescribed_cantilever

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def c1_mesh():
    return cantilever(20, 4, 4)


def het8_mesh():
    return cantilever(8, 2, 2, E=lognormal_E(7, 32, cov=0.3))


def khit_capacity(mesh, dmap, known=7):
    """row capacities as femKhit gives them (fem/fem_mloc.c:83-116): 24 per
    element touching the row's node"""
    n = int(dmap.max())
    cap = np.zeros(n, dtype=np.int64)
    for e in range(mesh.n_elems):
        for nd in mesh.enodes[e]:
            for d in range(3):
                r = dmap[int(nd) * known + d]
                if r > 0:
                    cap[r - 1] += 24
    return cap


@pytest.fixture
def policy_reset():
    yield
    G.set_mirror_policy(0)
    G.set_stop_rule(0)


def test_block_copy_refresh_and_policy(golden, policy_reset):
    """A node-major brick K goes to the device as 3x3 blocks.  Policy 0 sees
    host changes at every call; policy 1 keeps a valid copy until it is
    marked dirty; a pattern change rebuilds the copy."""
    mesh = c1_mesh()
    s = uo.OracleSystem(mesh)
    s.assemble(3)
    rl, c1, v = s.K_rowslot()
    K = G.RowSlotMatrix(s.n, rl, c1, v, capacity=rl + 3)
    b = G.Vec(s.F())
    ref = golden["c1_cg/u"]
    G.set_stop_rule(1)

    def solve():
        x = G.Vec(np.zeros(s.n))
        rv = G.femEqsCGwJ(K, b, x, 1e-12, 10 * s.n)
        assert rv == 0, G.last_error()
        return x.data

    u0 = solve()
    info = K.mirror_info()
    print("copy", info)
    assert info["bs"] == 3 and info["rows"] == s.n and info["nfill"] == int(rl.sum())
    assert rel(u0, ref) <= 1e-8
    # policy 0: host values scaled -> the next call sees them
    K.data *= 2.0
    assert rel(solve(), ref / 2) <= 1e-8
    assert K.mirror_info()["refresh_ms"] > 0
    # policy 1: the copy stays as it is until marked dirty
    G.set_mirror_policy(1)
    K.data *= 0.5
    stale = solve()
    assert K.mirror_info()["refresh_ms"] == 0.0
    assert rel(stale, ref / 2) <= 1e-8
    K.mark_dirty()
    assert rel(solve(), ref) <= 1e-8
    G.set_mirror_policy(0)
    # a new column (value 0) in row 1: the pattern changes, the blocks break
    f = int(K.frompos[0])
    K.pos[f + int(rl[0])] = s.n
    K.data[f + int(rl[0])] = 0.0
    assert rel(solve(), ref) <= 1e-8
    assert K.mirror_info()["bs"] == 1
    # femMatSetZeroBig clears host and device copies
    K.set_zero()
    assert not K.data.any()
    y = G.Vec(np.ones(s.n))
    assert G.femMatVecMultBig(K, G.Vec(np.ones(s.n)), y) == 0
    assert not y.data.any()
    K.release()


def test_scalar_copy_matches_spmv():
    """A matrix that is not made of node blocks (rows not a multiple of 3,
    ragged rows) goes to the device row by row; y = A x equals numpy's."""
    rng = np.random.default_rng(5)
    n = 301
    rows, cols, vals = [], [], []
    for i in range(n):
        k = int(rng.integers(1, 12))
        cs = np.unique(np.concatenate([[i], rng.integers(0, n, size=k)]))
        rng.shuffle(cs)
        rows.append(len(cs))
        cols.append(cs + 1)
        vals.append(rng.standard_normal(len(cs)))
    rl = np.array(rows)
    K = G.RowSlotMatrix(n, rl, np.concatenate(cols), np.concatenate(vals), capacity=rl + 2)
    x = rng.standard_normal(n)
    y = G.Vec(np.zeros(n))
    assert G.femMatVecMultBig(K, G.Vec(x), y) == 0
    A = np.zeros((n, n))
    for i in range(n):
        A[i, cols[i] - 1] = vals[i]
    assert np.abs(y.data - A @ x).max() <= 1e-13 * np.abs(A @ x).max()
    assert K.mirror_info()["bs"] == 1
    K.release()


def _check_K(Krs, golden, tag, which="K0"):
    ip, cs, vs = Krs.rows_csr()
    n = ip.size - 1
    assert int(golden[f"{tag}/n"]) == n
    assert np.array_equal(np.diff(ip), golden[f"{tag}/{which}_rowlen"])
    assert pattern_digest(ip, cs) == str(golden[f"{tag}/{which}_sha"])
    scale = np.abs(golden[f"{tag}/{which}_diag"]).max()
    w = probe_vector(n)
    # at most 81 entries per row, each within 1e-12 of max|K|, |w| <= 1
    assert np.abs(csr_matvec(ip, cs, vs, w) - golden[f"{tag}/{which}_Kw"]).max() <= 81 * 1e-12 * scale
    if f"{tag}/{which}_vals" in golden:
        assert np.array_equal(cs, golden[f"{tag}/{which}_cols"])
        assert np.abs(vs - golden[f"{tag}/{which}_vals"]).max() <= 1e-12 * scale


@pytest.mark.parametrize("tag,mk,fbc", [("c1_cg", c1_mesh, False), ("het8_cg", het8_mesh, False),
                                        ("het8_fbc", het8_mesh, True)])
def test_fill_K_into_rowslot(golden, tag, mk, fbc):
    """fem_fill_K(AF_NO) on the GPU into an empty row-slot K sized as femKhit
    sizes it: the reference's K0 (pattern digest, values); a refill after
    femMatSetZeroBig finds every column in place; a refill without it adds."""
    mesh = mk()
    dm = G.dof_map(mesh, fbc=fbc)
    if fbc:
        assert np.array_equal(dm, golden[f"{tag}/nDOFfld"])
    K = G.RowSlotMatrix.empty(int(dm.max()), khit_capacity(mesh, dm))
    assert G.fill_K(K, mesh, dm, 7, 3) == 0, G.last_error()
    print(tag, G.fill_K_timings())
    _check_K(K, golden, tag)
    K.set_zero()
    assert G.fill_K(K, mesh, dm, 7, 3) == 0
    _check_K(K, golden, tag)
    ip, cs, vs = K.rows_csr()
    assert G.fill_K(K, mesh, dm, 7, 3) == 0
    ip2, cs2, vs2 = K.rows_csr()
    assert np.array_equal(cs, cs2) and np.abs(vs2 - 2 * vs).max() <= 1e-12 * np.abs(vs).max()
    K.release()


def test_fill_K_row_overflow():
    """a row without room for its columns returns AF_ERR_BIG (the reference
    exits there, fem/fem_math.c:224-233)"""
    mesh = cantilever(3, 1, 1)
    dm = G.dof_map(mesh)
    cap = khit_capacity(mesh, dm)
    cap[5] = 10
    K = G.RowSlotMatrix.empty(int(dm.max()), cap)
    assert G.fill_K(K, mesh, dm, 7, 2) == -11
    assert "no free slot" in G.last_error()


@pytest.mark.parametrize("tag,mk", [("het8_fbc", het8_mesh)])
def test_fbc_fill_and_dropin_solve(golden, tag, mk, policy_reset):
    """-fbc end to end through the boundary: the GPU fills the compacted K
    (no support rows), the loads go to F, the drop-in CG solves with the
    reference's arguments: u equals the reference's -fbc u."""
    mesh = mk()
    dm = G.dof_map(mesh, fbc=True)
    n = int(dm.max())
    K = G.RowSlotMatrix.empty(n, khit_capacity(mesh, dm))
    assert G.fill_K(K, mesh, dm, 7, 3) == 0
    F = np.zeros(n)
    for nd, t, d, v in zip(mesh.nl_node, mesh.nl_type, mesh.nl_dir, mesh.nl_val):
        r = dm[int(nd) * 7 + int(d) - 1]
        if int(t) == 3 and r > 0:
            F[r - 1] += v
    assert np.abs(F - golden[f"{tag}/F1"]).max() == 0.0
    x = G.Vec(np.zeros(n))
    rv = G.femEqsCGwJ(K, G.Vec(F), x, 1e-21, n)
    info = G.last_solve()
    print("fbc dropin", rv, info, K.mirror_info())
    assert rv == int(golden[f"{tag}/rv"])
    assert rel(x.data, golden[f"{tag}/u"]) <= 1e-8
    # the x = 0 face loses whole nodes: the compacted K is still made of node blocks
    assert K.mirror_info()["bs"] == 3
    K.release()


def test_fill_K_prescribed_fbc_pattern(golden):
    """-fbc keeps supports of non-zero value in K: the GPU fill gives the
    reference's K0 for the 12x3x3 prescribed-displacement cantilever"""
    mesh = prescribed_cantilever(12, 3, 3, E=lognormal_E(13, 108, cov=0.3))
    dm = G.dof_map(mesh, fbc=True)
    assert np.array_equal(dm, golden["pd_fbc/nDOFfld"])
    K = G.RowSlotMatrix.empty(int(dm.max()), khit_capacity(mesh, dm))
    assert G.fill_K(K, mesh, dm, 7, 3) == 0
    _check_K(K, golden, "pd_fbc")
    K.release()
