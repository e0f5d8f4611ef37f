"""Result files of the native path against the compiled reference's own
femWriteRes output (fem/fem_io.c:648-734): an input with IDs (the result
file names the input's node and element IDs; golden het4_np.res) and the
-fbc form (zero supports left out of the DOF map and of u, their reactions
0; golden het8_fbc.res), on the reference's displacements; plus the native
-fbc solve against the reference's -fbc u."""

import os

import numpy as np
import pytest

from tests.resfile import close, read_res
from ufem_amd import gpu as G
from ufem_amd.mesh import cantilever, input_ids, lognormal_E, write_fem_input

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _compare(mine, ref):
    for k in ("header", "n_act", "n_dof_len", "n_nodes", "n_elems", "n_res"):
        assert mine[k] == ref[k], k
    for k in ("node_id", "dof_map", "r_node", "r_id", "r_dof", "elem_id", "res_from"):
        assert np.array_equal(mine[k], ref[k]), k
    for k in ("u", "R", "res"):
        assert close(mine[k], ref[k]), (k, np.abs(mine[k] - ref[k]).max())


def test_result_file_of_an_id_input(golden, tmp_path):
    mesh = cantilever(4, 2, 2, E=lognormal_E(11, 16, cov=0.6))
    ids = input_ids(mesh, 3)
    path = str(tmp_path / "np.fem")
    write_fem_input(mesh, path, ids=ids)
    m = G.read_model(path, preparsed=False)
    g = G.GpuSystem(m.mesh)
    g.set_ids(m.node_ids, m.elem_ids)
    g.assemble(3)
    g.set_u(golden["het4_np/u"])
    out = str(tmp_path / "np.res")
    g.write_results(out, sol_id=m.sol_id, dep_on=m.sol_dep)
    _compare(read_res(out), read_res(os.path.join(HERE, "golden", "het4_np.res")))


def test_result_file_fbc(golden, tmp_path):
    mesh = cantilever(8, 2, 2, E=lognormal_E(7, 32, cov=0.3))
    g = G.GpuSystem(mesh)
    g.set_fbc(True)
    dm, act = g.dof_map()
    assert act == int(golden["het8_fbc/n"]) and np.array_equal(dm, golden["het8_fbc/nDOFfld"])
    g.assemble(3)
    full = np.zeros(g.n)
    sel = dm.reshape(-1, 7)[:, :3].reshape(-1)
    full[sel > 0] = golden["het8_fbc/u"][sel[sel > 0] - 1]
    g.set_u(full)
    out = str(tmp_path / "fbc.res")
    g.write_results(out)
    _compare(read_res(out), read_res(os.path.join(HERE, "golden", "het8_fbc.res")))
    # the native solve: the supports are masked rows of the block matrix, and
    # the active DOFs of u are the reference's -fbc u
    info = g.solve("pcg", eps=1e-13)
    assert info["rv"] == 0
    u = g.u()[sel > 0]
    ref = golden["het8_fbc/u"]
    assert np.linalg.norm(u - ref) <= 1e-8 * np.linalg.norm(ref)
