"""The reference's own result file for the heterogeneous 4x2x2 cantilever
(tests/golden/het4_cg.res, written by femWriteRes of the compiled reference
through oracle/ref_harness.c) against the oracle: the element results of the
oracle's results pass on the reference's u, laid out as the brick's result
field, agree to the printed precision, and the u and reactions in the file
are the golden vectors."""

import os

import numpy as np

import ufem_oracle as uo
from tests.golden.make_golden import cases
from tests.resfile import close, read_res, reference_layout

HERE = os.path.dirname(os.path.abspath(__file__))


def test_reference_result_file_matches_the_oracle_results_pass(golden):
    r = read_res(os.path.join(HERE, "golden", "het4_cg.res"))
    mesh = {t: m for t, m, _ in cases()}["het4_cg"]
    assert r["header"] == (1, -1, 0.0)
    assert r["n_nodes"] == mesh.n_nodes and r["n_act"] == 3 * mesh.n_nodes
    assert np.array_equal(r["node_id"], np.arange(1, mesh.n_nodes + 1))
    dm = np.zeros((mesh.n_nodes, 7), dtype=np.int64)
    dm[:, :3] = np.arange(1, 3 * mesh.n_nodes + 1).reshape(-1, 3)
    assert np.array_equal(r["dof_map"], dm.reshape(-1))
    u = golden["het4_cg/u"]
    assert close(r["u"], u)
    assert close(r["R"], golden["het4_cg/R"])
    assert np.array_equal(r["r_node"], golden["het4_cg/Rnode"])
    assert np.array_equal(r["r_dof"], golden["het4_cg/Rdof"])
    assert r["n_elems"] == mesh.n_elems and r["n_res"] == 378 * mesh.n_elems
    assert np.array_equal(r["res_from"], 378 * np.arange(mesh.n_elems))
    s = uo.OracleSystem(mesh)
    s.assemble(3)
    s.set_u(u)
    expect = reference_layout(s.element_results(3))
    assert close(r["res"], expect)
    # the layout matters: without the cut at 378 values, or with epsilon
    # before sigma, the file does not match
    swapped = s.element_results(3)[:, :, list(range(6, 12)) + list(range(6))]
    assert not close(r["res"], reference_layout(swapped))
