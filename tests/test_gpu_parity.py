"""Parity of the HIP path with the reference (golden vectors written by the
compiled uFEM) and with the oracle restatement.

Bars: element matrices and assembled K within 1e-12 of max|K| (bit-exact
sparsity pattern), displacements within 1e-8 relative, reactions within 1e-6
relative (FP64 plain path)."""

import ctypes
import os

import numpy as np
import pytest

import ufem_oracle as uo
from tests.golden.make_golden import pattern_digest, probe_vector, csr_matvec, _diag_pos
from ufem_amd import gpu as G
from ufem_amd.mesh import cantilever, lognormal_E

pytestmark = pytest.mark.gpu

MESHES = {
    "c1": lambda: cantilever(20, 4, 4),
    "het8": lambda: cantilever(8, 2, 2, E=lognormal_E(7, 32, cov=0.3)),
    "het4": lambda: cantilever(4, 2, 2, E=lognormal_E(11, 16, cov=0.6)),
    "dist": lambda: cantilever(6, 3, 3, distort=0.25, seed=5, E=lognormal_E(3, 54, cov=0.3)),
    "cube": lambda: cantilever(1, 1, 1, L=1.0, B=1.0, H=1.0),
}


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oracle_tight(mesh, npts=3):
    s = uo.OracleSystem(mesh)
    s.assemble(npts)
    rv, it, rr = s.solve(0, 1e-15, 20 * s.n, 1)
    s.reactions(npts)
    return s


def test_device_present():
    info = G.device_info()
    print("device", info)
    assert info["n_cu"] > 0


@pytest.mark.parametrize("tag,mname,elems", [("c1_cg", "c1", (0, 319)), ("dist_cg", "dist", (0, 20, 53)),
                                             ("cube_cg", "cube", (0,))])
def test_element_stiffness_vs_reference(golden, tag, mname, elems):
    mesh = MESHES[mname]()
    xyz = np.stack([mesh.coords[mesh.enodes[e]] for e in elems])
    K3 = G.element_stiffness(xyz, mesh.E[list(elems)], mesh.nu[list(elems)], npts=3)
    K2 = G.element_stiffness(xyz, mesh.E[list(elems)], mesh.nu[list(elems)], npts=2)
    for i, e in enumerate(elems):
        ref = golden[f"{tag}/Ke_{e}"]
        scale = np.abs(ref).max()
        assert np.abs(K3[i] - ref).max() <= 1e-12 * scale        # 27 points, as the reference
        o2 = uo.element_stiffness(xyz[i], mesh.E[e], mesh.nu[e], 2)
        assert np.abs(K2[i] - o2).max() <= 1e-12 * scale         # 8 points vs oracle's 8 points
        if mname != "dist":                                      # boxes: 8 points are exact
            assert np.abs(K2[i] - ref).max() <= 1e-12 * scale


@pytest.mark.parametrize("tag,mname", [("c1_cg", "c1"), ("het8_cg", "het8"), ("het4_cg", "het4"),
                                       ("dist_cg", "dist"), ("cube_cg", "cube")])
def test_assembly_pattern_and_values(golden, tag, mname):
    mesh = MESHES[mname]()
    g = G.GpuSystem(mesh)
    g.fill_K(3)
    ip, cols, vals = g.K_csr()
    assert g.nnz() == mesh.expected_nnz()
    assert pattern_digest(ip, cols) == str(golden[f"{tag}/K0_sha"])      # bit-exact pattern
    scale = np.abs(golden[f"{tag}/K0_diag"]).max()
    w = probe_vector(g.n)
    assert np.abs(csr_matvec(ip, cols, vals, w) - golden[f"{tag}/K0_Kw"]).max() <= 1e-12 * scale * 81
    assert np.abs(vals[_diag_pos(ip, cols)] - golden[f"{tag}/K0_diag"]).max() <= 1e-12 * scale
    if f"{tag}/K0_vals" in golden:
        assert np.abs(vals - golden[f"{tag}/K0_vals"]).max() <= 1e-12 * scale
    g.add_loads()
    g.add_disps()
    ip1, cols1, vals1 = g.K_csr()
    assert pattern_digest(ip1, cols1) == str(golden[f"{tag}/K1_sha"])    # explicit zeros kept
    assert np.array_equal(g.F(), golden[f"{tag}/F1"])
    # supports: rows and columns of fixed DOFs zeroed, unit diagonal
    fixed = 3 * mesh.nl_node[mesh.nl_type == 1] + mesh.nl_dir[mesh.nl_type == 1] - 1
    rows = np.repeat(np.arange(g.n), np.diff(ip1))
    on_fixed = np.isin(rows, fixed) | np.isin(cols1, fixed)
    diag = rows == cols1
    assert np.all(vals1[on_fixed & ~diag] == 0.0)
    assert np.all(vals1[on_fixed & diag & np.isin(rows, fixed)] == 1.0)
    keep = ~on_fixed
    assert np.abs(vals1[keep] - vals[keep]).max() == 0.0


@pytest.mark.parametrize("tag,mname", [("c1_cg", "c1"), ("het8_cg", "het8"), ("dist_cg", "dist"),
                                       ("cube_cg", "cube")])
def test_solve_matches_reference(golden, tag, mname):
    mesh = MESHES[mname]()
    g = G.GpuSystem(mesh)
    g.assemble(3)
    info = g.solve("pcg", eps=1e-14, rule=1)
    assert info["rv"] == 0, info
    u = g.u()
    err_u = rel(u, golden[f"{tag}/u"])
    R, rn, rd = g.reactions()
    err_R = rel(R, golden[f"{tag}/R"])
    print(tag, "iters", info["iters"], "u rel err", err_u, "R rel err", err_R, "true relres", info["true_relres"])
    assert err_u <= 1e-8
    assert err_R <= 1e-6
    assert np.array_equal(rn, golden[f"{tag}/Rnode"]) and np.array_equal(rd, golden[f"{tag}/Rdof"])
    # equilibrium: reactions balance the applied load
    Rz = R[rd == 3].sum()
    assert abs(Rz + mesh.nl_val[mesh.nl_type == 3].sum()) <= 1e-8 * abs(mesh.nl_val[mesh.nl_type == 3].sum())


def test_unconverged_reference_case_vs_tight_oracle(golden):
    """het4: the reference stops at maxIt = n unconverged (AF_ERR); the GPU
    converges, and matches the oracle solved to 1e-15."""
    mesh = MESHES["het4"]()
    assert int(golden["het4_cg/rv"]) == -1
    o = oracle_tight(mesh)
    g = G.GpuSystem(mesh)
    g.assemble(3)
    info = g.solve("pcg", eps=1e-14, rule=1, maxit=10 * mesh.n_dof)
    assert info["rv"] == 0
    assert rel(g.u(), o.u()) <= 1e-8
    assert rel(g.reactions()[0], o.R()[0]) <= 1e-6
    # with the reference's own arguments the GPU also stops unconverged
    g2 = G.GpuSystem(mesh)
    g2.assemble(3)
    info2 = g2.solve("pcg", eps=1e-21, rule=0, maxit=0)
    assert info2["rv"] == -1 and info2["iters"] == g2.n


def test_medium_heterogeneous_vs_oracle():
    mesh = cantilever(40, 8, 8, E=lognormal_E(99, 40 * 64, cov=0.3))
    o = oracle_tight(mesh, npts=2)
    g = G.GpuSystem(mesh)
    g.assemble(2)
    info = g.solve("pcg", eps=1e-13, rule=1)
    assert info["rv"] == 0
    e_u, e_R = rel(g.u(), o.u()), rel(g.reactions()[0], o.R()[0])
    print("medium iters", info["iters"], "u", e_u, "R", e_R)
    assert e_u <= 1e-8 and e_R <= 1e-6


def test_dropin_cgwj_reference_arguments(golden):
    """femEqsCGwJ through the C-ABI on uFEM's own row-slot K (after supports)
    with the arguments femSolve passes (eps 1e-21, maxIt = n)."""
    mesh = MESHES["c1"]()
    s = uo.OracleSystem(mesh)
    s.assemble(3)
    rl, c1, v = s.K_rowslot()
    K = G.RowSlotMatrix(s.n, rl, c1, v, capacity=rl + 3)
    b = G.Vec(s.F())
    x = G.Vec(np.zeros(s.n))
    G.set_stop_rule(0)
    rv = G.femEqsCGwJ(K, b, x, 1e-21, s.n)
    info = G.last_solve()
    print("dropin", rv, info)
    assert rv == 0
    assert rel(x.data, golden["c1_cg/u"]) <= 1e-8
    assert abs(info["iters"] - (int(golden["c1_cg/spmv_calls"]) - 1)) <= 20


def test_dropin_bicgstab(golden):
    mesh = MESHES["c1"]()
    s = uo.OracleSystem(mesh)
    s.assemble(3)
    rl, c1, v = s.K_rowslot()
    K = G.RowSlotMatrix(s.n, rl, c1, v)
    b = G.Vec(s.F())
    x = G.Vec(np.zeros(s.n))
    G.set_stop_rule(1)
    # the stop is confirmed on the FP64 true residual, whose rounding floor on
    # this beam lies near 1e-12: 1e-11 is the tightest reliable tolerance
    rv = G.femEqsBiCCSwJ(K, b, x, 1e-11, s.n)
    G.set_stop_rule(0)
    assert rv == 0, G.last_solve()
    assert rel(x.data, golden["c1_cg/u"]) <= 1e-8


def test_native_bicgstab(golden):
    mesh = MESHES["het8"]()
    g = G.GpuSystem(mesh)
    g.assemble(3)
    info = g.solve("bicgstab", eps=1e-11, rule=1)
    assert info["rv"] == 0, info
    assert rel(g.u(), golden["het8_cg/u"]) <= 1e-8


@pytest.mark.parametrize("mname,npts", [("c1", 3), ("c1", 2), ("dist", 3), ("het8", 3)])
def test_element_results_vs_oracle(mname, npts):
    """Strains and stresses at the Gauss points (results pass of
    fem_fill_K(AF_YES)): on the oracle's displacements they agree with the
    oracle to 1e-12 of the largest value; from the GPU's own solve, to the
    displacement parity."""
    mesh = MESHES[mname]()
    o = oracle_tight(mesh, npts=npts)
    ref = o.element_results(npts)
    g = G.GpuSystem(mesh)
    g.assemble(npts)
    g.set_u(o.u())
    res = g.element_results(npts)
    for sl in (slice(0, 6), slice(6, 12)):
        scale = np.abs(ref[:, :, sl]).max()
        assert np.abs(res[:, :, sl] - ref[:, :, sl]).max() <= 1e-12 * scale
    info = g.solve("pcg", eps=1e-14, rule=1)
    assert info["rv"] == 0
    own = g.element_results(npts)
    assert rel(own, ref) <= 1e-8


def test_result_file_matches_reference(golden, tmp_path):
    """ufem_sys_write_results on the reference's u for the heterogeneous 4x2x2
    cantilever against the reference's own femWriteRes file: the same layout
    and integers, and every number equal to the 7 digits printed."""
    from tests.resfile import close, read_res
    mesh = MESHES["het4"]()
    g = G.GpuSystem(mesh)
    g.assemble(3)
    g.set_u(golden["het4_cg/u"])
    path = str(tmp_path / "het4.res")
    g.write_results(path)
    mine = read_res(path)
    ref = read_res(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "het4_cg.res"))
    for k in ("header", "n_act", "n_dof_len", "n_nodes", "n_elems", "n_res"):
        assert mine[k] == ref[k], k
    for k in ("node_id", "dof_map", "r_node", "r_id", "r_dof", "elem_id", "res_from"):
        assert np.array_equal(mine[k], ref[k]), k
    for k in ("u", "R", "res"):
        assert close(mine[k], ref[k]), (k, np.abs(mine[k] - ref[k]).max())


def test_prescribed_displacements_vs_eliminated_system(golden, tmp_path):
    """Non-zero prescribed displacements (some listed before the supports): the
    loads of the free rows lose K(:, P) v in one SpMV and the supports zero
    full rows and columns, so u solves the eliminated system (a dense solve
    here) to 1e-8; the reactions are the support forces (K u - F) at the
    prescribed DOFs, within 1e-6, and balance the applied load."""
    import scipy.sparse as sp
    from ufem_amd.mesh import prescribed_cantilever
    mesh = prescribed_cantilever(12, 3, 3, E=lognormal_E(13, 108, cov=0.3))
    o = uo.OracleSystem(mesh)
    o.fill_K(3)
    o.add_loads()
    ip, c, v = o.K_csr()
    K0 = sp.csr_matrix((v, c, ip), shape=(o.n, o.n)).toarray()
    F0 = o.F()
    pres = {3 * n + d - 1: val for n, t, d, val in zip(mesh.nl_node, mesh.nl_type, mesh.nl_dir, mesh.nl_val) if t == 1}
    P = np.array(sorted(pres))
    vals = np.array([pres[p] for p in P])
    free = np.setdiff1d(np.arange(o.n), P)
    u = np.zeros(o.n)
    u[P] = vals
    u[free] = np.linalg.solve(K0[np.ix_(free, free)], F0[free] - K0[np.ix_(free, P)] @ vals)
    g = G.GpuSystem(mesh)
    g.assemble(3)
    for solver in ("pcg", "ssor"):
        info = g.solve(solver, eps=1e-13, rule=1)
        assert info["rv"] == 0, (solver, info)
        ug = g.u()
        assert rel(ug, u) <= 1e-8, solver
        assert np.abs(ug[P] - vals).max() <= 1e-12 * np.abs(vals).max()
    R, rn, rd = g.reactions()
    Rd = (K0 @ u - F0)[3 * rn + rd - 1]
    assert rel(R, Rd) <= 1e-6
    load = mesh.nl_val[mesh.nl_type == 3].sum()
    assert abs(R[rd == 3].sum() + load) <= 1e-6 * np.abs(R[rd == 3]).sum()
    # the result file lists the zero-valued supports only, as the reference's
    # list (fem/fem_sol.c:676-722) that the pd golden case recorded
    from tests.resfile import read_res
    g.write_results(str(tmp_path / "pd.res"))
    res = read_res(str(tmp_path / "pd.res"))
    assert np.array_equal(res["r_node"], golden["pd_cg/Rnode"])
    assert np.array_equal(res["r_dof"], golden["pd_cg/Rdof"])
    zero = np.array([mesh.nl_val[i] == 0.0 for i in np.nonzero(mesh.nl_type == 1)[0]])
    assert rel(res["R"], R[zero]) <= 1e-6


def oracle_ssor(mesh, eps, npts=3):
    s = uo.OracleSystem(mesh)
    s.assemble(npts)
    rv, it, rr = s.solve(3, eps, 20 * s.n, 1)
    return s, rv, it


@pytest.mark.parametrize("mname", ["c1", "het8", "het4", "dist"])
def test_native_ssor_vs_oracle(golden, mname):
    """CG with the SSOR preconditioner on the device (colour-ordered sweeps)
    against the oracle's natural-order femEqsCGwSSOR: the same displacements
    and reactions to the parity bars; fewer iterations than Jacobi PCG."""
    mesh = MESHES[mname]()
    o, rv_o, it_o = oracle_ssor(mesh, 1e-13)
    assert rv_o == 0
    g = G.GpuSystem(mesh)
    g.assemble(3)
    info = g.solve("ssor", eps=1e-13, rule=1)
    jac = g.solve("pcg", eps=1e-13, rule=1)
    g.solve("ssor", eps=1e-13, rule=1)
    print(mname, "ssor", info["iters"], "colours", info["outer"], "oracle ssor", it_o, "jacobi", jac["iters"])
    assert info["rv"] == 0 and info["true_relres"] <= 1e-10     # the recurrence drifts below ~1e-12
    assert rel(g.u(), o.u()) <= 1e-8
    if mname != "het4":                  # the reference stops unconverged on het4
        assert rel(g.u(), golden[f"{mname}_cg/u"]) <= 1e-8
    o.reactions(3)
    assert rel(g.reactions()[0], o.R()[0]) <= 1e-6
    assert info["iters"] < jac["iters"]
    assert info["outer"] >= 2


def test_medium_ssor_vs_oracle():
    mesh = cantilever(40, 8, 8, E=lognormal_E(99, 40 * 64, cov=0.3))
    o, rv_o, it_o = oracle_ssor(mesh, 1e-12, npts=2)
    g = G.GpuSystem(mesh)
    g.assemble(2)
    info = g.solve("ssor", eps=1e-12, rule=1)
    print("medium ssor", info["iters"], "colours", info["outer"], "oracle", it_o)
    assert info["rv"] == 0
    assert rel(g.u(), o.u()) <= 1e-8
    assert info["iters"] <= 1.5 * it_o


def test_dropin_ssor_reference_arguments(golden):
    """femEqsCGwSSOR through the C-ABI on uFEM's own row-slot K (scalar rows,
    so the colouring is of DOFs), the reference rule, eps 1e-21, maxIt = n;
    a second call on the same matrix reuses the colour order."""
    mesh = MESHES["c1"]()
    s = uo.OracleSystem(mesh)
    s.assemble(3)
    rl, c1, v = s.K_rowslot()
    K = G.RowSlotMatrix(s.n, rl, c1, v, capacity=rl + 3)
    b = G.Vec(s.F())
    G.set_stop_rule(0)
    for _ in range(2):
        x = G.Vec(np.zeros(s.n))
        rv = G.femEqsCGwSSOR(K, b, x, 1e-21, s.n)
        info = G.last_solve()
        print("dropin ssor", rv, info)
        assert rel(x.data, golden["c1_cg/u"]) <= 1e-8
    G.set_stop_rule(1)
    x = G.Vec(np.zeros(s.n))
    rv = G.femEqsCGwSSOR(K, b, x, 1e-12, s.n)
    G.set_stop_rule(0)
    assert rv == 0 and G.last_solve()["true_relres"] <= 1e-10
    assert rel(x.data, golden["c1_cg/u"]) <= 1e-8


def test_dropin_spmv_and_reductions():
    mesh = MESHES["het8"]()
    s = uo.OracleSystem(mesh)
    s.assemble(3)
    rl, c1, v = s.K_rowslot()
    K = G.RowSlotMatrix(s.n, rl, c1, v)
    x = G.Vec(np.sin(np.arange(s.n)))
    y = G.Vec(np.zeros(s.n))
    L = G.lib()
    assert L.femMatVecMultBig(K.s, x.s, y.s) == 0
    ref = s.spmv(x.data)
    assert np.abs(y.data - ref).max() <= 1e-13 * np.abs(ref).max()
    assert abs(L.femVecVecMultBig(x.s, y.s) - float(x.data @ ref)) <= 1e-12 * abs(float(x.data @ ref))
    assert abs(L.femVecNormBig(x.s) - np.linalg.norm(x.data)) <= 1e-14 * np.linalg.norm(x.data)
    assert L.femMatNormBig(K.s) == pytest.approx(s.mat_norm(), rel=1e-15)
    z = G.Vec(np.zeros(s.n))
    assert L.femVecLinCombBig(2.0, x.s, -3.0, y.s, z.s) == 0
    assert np.abs(z.data - (2 * x.data - 3 * y.data)).max() <= 1e-15 * np.abs(z.data).max()


@pytest.mark.parametrize("ex", ["persistent", "fused", "classic"])
def test_exec_forms_agree(ex):
    mesh = cantilever(30, 6, 6, E=lognormal_E(5, 30 * 36, cov=0.3))
    a = G.GpuSystem(mesh)
    a.assemble(2)
    ia = a.solve("pcg", eps=1e-10, rule=1, exec="streamed")
    b = G.GpuSystem(mesh)
    b.assemble(2)
    ib = b.solve("pcg", eps=1e-10, rule=1, exec=ex)
    print("streamed", ia, ex, ib)
    assert ia["rv"] == 0 and ib["rv"] == 0
    assert abs(ia["iters"] - ib["iters"]) <= 2
    assert rel(a.u(), b.u()) <= 1e-9


def test_bitwise_reproducible():
    mesh = cantilever(24, 6, 6, E=lognormal_E(8, 24 * 36, cov=0.3))
    us = []
    for _ in range(2):
        g = G.GpuSystem(mesh)
        g.assemble(2)
        g.solve("pcg", eps=1e-10, rule=1)
        us.append(g.u())
    assert np.array_equal(us[0], us[1])


def test_mixed_precision_error_vs_fp64(golden):
    mesh = MESHES["c1"]()
    g = G.GpuSystem(mesh)
    g.assemble(3)
    info = g.solve("mixed", eps=1e-10, rule=1, maxit=5000)
    err = rel(g.u(), golden["c1_cg/u"])
    print("mixed", info, "u rel err vs FP64 reference", err)
    assert info["rv"] == 0
    assert info["true_relres"] <= 1e-10
    assert err <= 1e-7


def test_error_paths():
    # inverted element -> AF_ERR_VAL from fill_K
    mesh = cantilever(2, 1, 1)
    bad = mesh.enodes.copy()
    bad[0] = bad[0][[4, 5, 6, 7, 0, 1, 2, 3]]
    mesh.enodes = bad
    g = G.GpuSystem(mesh)
    with pytest.raises(RuntimeError, match="Jacobian"):
        g.fill_K(2)
    # too few iterations -> AF_ERR, the iterate is still returned
    mesh = MESHES["c1"]()
    g = G.GpuSystem(mesh)
    g.assemble(2)
    info = g.solve("pcg", eps=1e-12, rule=1, maxit=5)
    assert info["rv"] == -1 and info["iters"] == 5
    # zero load -> u = 0, AF_OK
    mesh = MESHES["c1"]()
    mesh.nl_val = np.where(mesh.nl_type == 3, 0.0, mesh.nl_val)
    g = G.GpuSystem(mesh)
    g.assemble(2)
    info = g.solve("pcg", eps=1e-12, rule=1)
    assert info["rv"] == 0 and np.all(g.u() == 0.0)
    # zero diagonal through the drop-in -> AF_ERR_ZER
    K = G.RowSlotMatrix(3, np.array([1, 1, 1]), np.array([1, 2, 2]), np.array([1.0, 0.0, 1.0]))
    rv = G.femEqsCGwJ(K, G.Vec(np.ones(3)), G.Vec(np.zeros(3)), 1e-10, 10)
    assert rv == G.AF_ERR_ZER


def test_timing_modes_leave_results_intact():
    """the kernel timers (HIP events on the system's stream): every mode runs
    and reports its byte count, the stream read moves the stored matrix, an
    unknown mode is refused, and K and the solve are unchanged afterwards"""
    mesh = MESHES["c1"]()
    g = G.GpuSystem(mesh)
    g.assemble(2)
    ref = g.solve("pcg", eps=1e-12, rule=1)
    u0 = g.u()
    ip, cols, vals = g.K_csr()
    nnzb = len(vals) // 9
    alg = nnzb * 76.0 + 16.0 * g.n
    for mode in ("spmv", "spmv32", "pcg_q", "cg1_w", "copy", "read", "read_nt"):
        ms, by = g.time_kernel(mode, reps=3)
        assert ms > 0.0 and by > 0.0, mode
        if mode in ("spmv", "pcg_q", "cg1_w"):
            assert by == alg
        if mode.startswith("read"):
            # stored values and columns: the real blocks plus the slice padding
            assert nnzb * 76.0 <= by <= 1.5 * nnzb * 76.0
    ms, by = ctypes.c_double(0.0), ctypes.c_double(0.0)
    assert G.lib().ufem_sys_time_spmv(g.h, 1, 8, ctypes.byref(ms), ctypes.byref(by)) == G.AF_ERR_VAL
    assert "unknown kernel mode" in G.last_error()
    ip2, cols2, vals2 = g.K_csr()
    assert np.array_equal(ip, ip2) and np.array_equal(cols, cols2) and np.array_equal(vals, vals2)
    again = g.solve("pcg", eps=1e-12, rule=1)
    assert again["iters"] == ref["iters"] and np.array_equal(g.u(), u0)
    g.close()
