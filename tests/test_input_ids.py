"""The library's reader of uFEM's input file (host code, no GPU): the form
with IDs (the fem binary without -p) is renumbered to the positions the
compiled reference's femRenumIndex gives (golden het4_np: element nodes,
load nodes, element material sets); the preparsed form of the same mesh
reads to the same model; a repeated ID resolves to its first item and an ID
no item carries is AF_ERR_VAL, as femRenumIndex does (fem/fem_io.c:50-80)."""

import numpy as np

from ufem_amd import gpu as G
from ufem_amd.mesh import cantilever, input_ids, lognormal_E, write_fem_input


def het4():
    return cantilever(4, 2, 2, E=lognormal_E(11, 16, cov=0.6))


def test_ids_renumbered_as_the_reference(golden, tmp_path):
    mesh = het4()
    ids = input_ids(mesh, 3)
    path = str(tmp_path / "np.fem")
    write_fem_input(mesh, path, ids=ids)
    m = G.read_model(path, preparsed=False)
    assert np.array_equal(m.mesh.enodes.ravel(), golden["het4_np/eNodesL"])
    assert np.array_equal(m.mesh.nl_node, golden["het4_np/nlNode"])
    assert np.array_equal(m.elem_mat, golden["het4_np/eMP"])
    assert np.array_equal(m.node_ids, ids[0]) and np.array_equal(m.elem_ids, ids[1])
    # the renumbered model is the generated mesh
    assert np.array_equal(m.mesh.enodes, mesh.enodes)
    assert np.array_equal(m.mesh.coords, mesh.coords)
    assert np.array_equal(m.mesh.E, mesh.E) and np.array_equal(m.mesh.nl_val, mesh.nl_val)
    p2 = str(tmp_path / "p.fem")
    write_fem_input(mesh, p2)
    mp = G.read_model(p2, preparsed=True)
    assert np.array_equal(mp.mesh.enodes, m.mesh.enodes) and np.array_equal(mp.mesh.nl_node, m.mesh.nl_node)


def test_repeated_and_missing_ids(tmp_path):
    mesh = cantilever(2, 1, 1)
    nid, eid, mid = input_ids(mesh, 1)
    # node 5 takes node 2's ID: every reference to that ID means node 2 (the first)
    nid2 = nid.copy()
    nid2[5] = nid[2]
    path = str(tmp_path / "dup.fem")
    write_fem_input(mesh, path, ids=(nid2, eid, mid))
    m = G.read_model(path, preparsed=False)
    want = mesh.enodes.copy()
    want[want == 5] = 2
    assert np.array_equal(m.mesh.enodes, want)
    # an element naming an ID that no node has
    write_fem_input(mesh, path, ids=(nid, eid, mid))
    lines = open(path).read().split("\n")
    k = next(i for i, ln in enumerate(lines) if ln.strip() == str(8 * mesh.n_elems))
    row = lines[k + 1].split()
    row[0] = "999999"
    lines[k + 1] = " ".join(row)
    open(path, "w").write("\n".join(lines))
    rv, msg = G.model_error(path, preparsed=False)
    assert rv == -3 and "999999" in msg


def test_large_id_input_is_linear(tmp_path):
    """a 60x12x12 mesh with IDs (8,640 elements, 10,309 nodes) reads in well
    under a second (the linear search would make ~9e8 comparisons)"""
    import time
    mesh = cantilever(60, 12, 12)
    path = str(tmp_path / "big.fem")
    write_fem_input(mesh, path, ids=input_ids(mesh, 5))
    t0 = time.perf_counter()
    m = G.read_model(path, preparsed=False)
    dt = time.perf_counter() - t0
    assert np.array_equal(m.mesh.enodes, mesh.enodes)
    assert dt < 2.0, dt
