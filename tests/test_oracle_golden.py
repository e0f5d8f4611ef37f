"""The oracle (clean-room C restatement) against golden vectors written by the
compiled uFEM reference (tests/golden/make_golden.py).  Serial mode must be
bit-for-bit equal: K before/after supports, F, u, reactions, K_e, ||A||, the
CG/BiCGStab iteration counts and return codes."""

import numpy as np
import pytest

import ufem_oracle as uo
from ufem_amd.mesh import cantilever, lognormal_E, prescribed_cantilever
from tests.golden.make_golden import pattern_digest, probe_vector, csr_matvec, _diag_pos

MESHES = {
    "c1": lambda: cantilever(20, 4, 4),
    "het8": lambda: cantilever(8, 2, 2, E=lognormal_E(7, 32, cov=0.3)),
    "het4": lambda: cantilever(4, 2, 2, E=lognormal_E(11, 16, cov=0.6)),
    "dist": lambda: cantilever(6, 3, 3, distort=0.25, seed=5, E=lognormal_E(3, 54, cov=0.3)),
    "cube": lambda: cantilever(1, 1, 1, L=1.0, B=1.0, H=1.0),
    "pd": lambda: prescribed_cantilever(12, 3, 3, E=lognormal_E(13, 108, cov=0.3)),
}

# (tag, mesh, solver id, eps)
CASES = [
    ("c1_cg", "c1", 0, 1e-21),
    ("c1_cg_eps12", "c1", 0, 1e-12),
    ("c1_bicgs", "c1", 1, 1e-21),
    ("het8_cg", "het8", 0, 1e-21),
    ("het8_bicgs", "het8", 1, 1e-21),
    ("het4_cg", "het4", 0, 1e-21),
    ("dist_cg", "dist", 0, 1e-21),
    ("cube_cg", "cube", 0, 1e-21),
    ("pd_cg", "pd", 0, 1e-12),       # prescribed non-zero displacements, some before the supports
]


@pytest.mark.parametrize("tag,mname,solver,eps", CASES)
def test_oracle_bitwise_vs_reference(golden, tag, mname, solver, eps):
    mesh = MESHES[mname]()
    assert np.array_equal(mesh.E, golden[f"{tag}/E"]), "mesh generator drifted"
    s = uo.OracleSystem(mesh, threads=1)
    assert s.n == int(golden[f"{tag}/n"])
    s.fill_K(3)
    w = probe_vector(s.n)
    for kname in ("K0", "K1"):
        if kname == "K1":
            s.add_loads()
            s.add_disps()
            assert np.array_equal(s.F(), golden[f"{tag}/F1"])
        rl, c1, v = s.K_rowslot()
        ip, cs, vs = uo.rowslot_to_csr(rl, c1, v)
        assert np.array_equal(rl, golden[f"{tag}/{kname}_rowlen"])
        assert pattern_digest(ip, cs) == str(golden[f"{tag}/{kname}_sha"])
        k = int(rl[:12].sum())
        assert np.array_equal(c1[:k], golden[f"{tag}/{kname}_head_cols"])   # insertion order
        assert np.array_equal(v[:k], golden[f"{tag}/{kname}_head_vals"])
        assert np.array_equal(vs[_diag_pos(ip, cs)], golden[f"{tag}/{kname}_diag"])
        assert np.array_equal(csr_matvec(ip, cs, vs, w), golden[f"{tag}/{kname}_Kw"])
        if f"{tag}/{kname}_vals" in golden:
            assert np.array_equal(vs, golden[f"{tag}/{kname}_vals"])
    assert s.mat_norm() == float(golden[f"{tag}/normA"])
    rv, iters, _ = s.solve(solver=solver, eps=eps, maxit=-1, rule=0)
    assert rv == int(golden[f"{tag}/rv"])
    calls = int(golden[f"{tag}/spmv_calls"])
    if solver == 0:
        assert iters + 1 == calls            # one SpMV for r0, one per iteration
    assert np.array_equal(s.u(), golden[f"{tag}/u"])
    s.reactions(3)
    R, rn, rd = s.R()
    assert np.array_equal(rn, golden[f"{tag}/Rnode"])
    assert np.array_equal(rd, golden[f"{tag}/Rdof"])
    assert np.array_equal(R, golden[f"{tag}/R"])


@pytest.mark.parametrize("tag,mname,elems", [("c1_cg", "c1", (0, 319)), ("dist_cg", "dist", (0, 20, 53)),
                                             ("cube_cg", "cube", (0,))])
def test_element_stiffness_bitwise(golden, tag, mname, elems):
    mesh = MESHES[mname]()
    for e in elems:
        Ke = uo.element_stiffness(mesh.coords[mesh.enodes[e]], mesh.E[e], mesh.nu[e], 3)
        assert np.array_equal(Ke, golden[f"{tag}/Ke_{e}"])


def test_threaded_reference_run_close(golden):
    """The reference with -t 4 assembles with per-row mutexes, so its column
    insertion and summation order depend on thread timing (the threaded run
    is not reproducible run to run); the oracle with 4 threads agrees to
    rounding and within a few iterations."""
    mesh = MESHES["c1"]()
    s = uo.OracleSystem(mesh, threads=4)
    s.assemble(3)
    rv, iters, _ = s.solve(0, 1e-21, -1, 0)
    assert rv == int(golden["c1_cg_t4/rv"])
    assert abs(iters + 1 - int(golden["c1_cg_t4/spmv_calls"])) <= 3
    ref = golden["c1_cg_t4/u"]
    assert np.linalg.norm(s.u() - ref) <= 1e-12 * np.linalg.norm(ref)


def test_reference_double_counts_prescribed_couplings_outside_the_window(golden):
    """With non-zero prescribed displacements the reference subtracts K(:, p) v
    from F for every row (fem/fem_loads.c:210-224) but zeroes column p only in
    its row window (fem/fem_math.c:592-620), so rows outside the window count
    the coupling twice: the reference's u (reproduced bitwise by the oracle,
    test above) is far from the solution of the eliminated system, which the
    GPU path solves."""
    import scipy.sparse as sp
    mesh = MESHES["pd"]()
    s = uo.OracleSystem(mesh)
    s.fill_K(3)
    s.add_loads()
    ip, c, v = s.K_csr()
    K0 = sp.csr_matrix((v, c, ip), shape=(s.n, s.n)).toarray()
    F0 = s.F()
    pres = {3 * n + d - 1: val for n, t, d, val in zip(mesh.nl_node, mesh.nl_type, mesh.nl_dir, mesh.nl_val) if t == 1}
    P = np.array(sorted(pres))
    vals = np.array([pres[p] for p in P])
    free = np.setdiff1d(np.arange(s.n), P)
    u = np.zeros(s.n)
    u[P] = vals
    u[free] = np.linalg.solve(K0[np.ix_(free, free)], F0[free] - K0[np.ix_(free, P)] @ vals)
    ref = golden["pd_cg/u"]
    assert np.abs(ref[P] - vals).max() <= 1e-12 * np.abs(vals).max()
    assert np.linalg.norm(ref - u) > 1.0 * np.linalg.norm(u)
