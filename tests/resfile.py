"""Parser of the reference's result file (femWriteRes, fem/fem_io.c:648-734),
for the tests."""

import numpy as np


def read_res(path):
    with open(path) as f:
        tok = [ln.split() for ln in f.read().splitlines()]
    it = iter(tok)
    out = {}
    assert next(it) == ["S"]
    h = next(it)
    out["header"] = (int(h[0]), int(h[1]), float(h[2]))
    n_act, n_len, nn = (int(next(it)[0]) for _ in range(3))
    out["n_act"], out["n_dof_len"], out["n_nodes"] = n_act, n_len, nn
    out["node_id"] = np.array([int(next(it)[0]) for _ in range(nn)])
    out["dof_map"] = np.array([int(next(it)[0]) for _ in range(n_len)])
    nu = int(next(it)[0])
    out["u"] = np.array([float(next(it)[0]) for _ in range(nu)])
    nr = int(next(it)[0])
    rows = [next(it) for _ in range(nr)]
    out["r_node"] = np.array([int(r[0]) for r in rows])
    out["r_id"] = np.array([int(r[1]) for r in rows])
    out["r_dof"] = np.array([int(r[2]) for r in rows])
    out["R"] = np.array([float(r[3]) for r in rows])
    ne, nres = (int(v) for v in next(it))
    out["n_elems"], out["n_res"] = ne, nres
    out["elem_id"] = np.array([int(next(it)[0]) for _ in range(ne)])
    out["res_from"] = np.array([int(next(it)[0]) for _ in range(ne)])
    out["res"] = np.array([float(next(it)[0]) for _ in range(nres)])
    assert next(it, None) is None
    return out


def reference_layout(gp_results):
    """[ne, 27, 12] -> the element result field of a brick: 20 repeating values
    per point (sigma, epsilon, 8 zeros for a linear material) cut at 14 x 27
    values per element"""
    ne = gp_results.shape[0]
    full = np.zeros((ne, 27, 20))
    full[:, :, :12] = gp_results
    return full.reshape(ne, -1)[:, : 14 * 27].reshape(-1)


def close(a, b, rel=2e-6):
    """equal to the 7 significant digits %e prints"""
    return np.all(np.abs(a - b) <= rel * np.abs(b) + 1e-300)
