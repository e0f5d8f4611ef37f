"""Synthetic cantilever meshes of 8-node bricks (uFEM element type 9).

Geometry, numbering and load conventions follow the reference's preparsed
input (fem/feminput.5) and the brick's node order (fem/fem_e009.c:36-69):

* box [0, L] x [0, B] x [0, H], nx * ny * nz bricks;
* node position i + (nx+1) (j + (ny+1) k) -- x fastest, so uFEM's node-major
  DOF numbering (fem/fem_sol.c:1297-1350) puts DOF 3 p + d at node p;
* element e = i + nx (j + ny k) with corners (-,-,-) (+,-,-) (+,+,-) (-,+,-)
  (-,-,+) (+,-,+) (+,+,+) (-,+,+) in (xi, eta, zeta) = (x, y, z);
* supports: all three displacements zero on every x = 0 node (type 1);
* load: total Fz = -10 kN shared equally by the x = L nodes (type 3).

Per-element Young's moduli come from a counter-based generator (splitmix64 of
seed and element id, Box-Muller), so the same field is reproducible anywhere.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.fft as sfft

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def counter_normal(seed: int, ids: np.ndarray, stream: int = 0) -> np.ndarray:
    """Standard normals indexed by (seed, stream, id): Box-Muller on two
    splitmix64 draws, so element e always receives the same value."""
    ids = np.asarray(ids, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = (np.uint64(seed) * np.uint64(0x632BE59BD9B4E019)
                + np.uint64(stream) * np.uint64(0x9E3779B97F4A7C15)) & _M64
        k = (base + ids * np.uint64(2)) & _M64
    a = _splitmix64(k)
    b = _splitmix64(k + np.uint64(1))
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)
    u2 = (b >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def lognormal_E(seed: int, n_elems: int, mean: float = 30e9, cov: float = 0.3,
                stream: int = 0) -> np.ndarray:
    """i.i.d. lognormal moduli with the given mean and coefficient of variation."""
    s2 = math.log1p(cov * cov)
    mu = math.log(mean) - 0.5 * s2
    z = counter_normal(seed, np.arange(n_elems), stream)
    return np.exp(mu + math.sqrt(s2) * z)


def grf_E(seed: int, nx: int, ny: int, nz: int, h: tuple[float, float, float],
          corr_len: float = 0.2, log_std: float = 0.5, mean: float = 30e9,
          stream: int = 1) -> np.ndarray:
    """E = exp(mu + log_std * g) with g a unit-variance Gaussian random field on
    the element centres.  Covariance exp(-r^2 / corr_len^2) (squared
    exponential): white noise on a grid padded by 4 correlation lengths is
    convolved, by FFT, with the kernel exp(-2 r^2 / corr_len^2) normalised to
    unit l2 norm, whose self-convolution is that covariance.  mu =
    ln(mean) - log_std^2 / 2 so that E has the requested mean."""
    pad = [int(math.ceil(4.0 * corr_len / hi)) for hi in h]
    shape = (nz + 2 * pad[2], ny + 2 * pad[1], nx + 2 * pad[0])
    w = counter_normal(seed, np.arange(int(np.prod(shape))), stream).reshape(shape)
    axes = []
    for n_, hi in zip(shape, (h[2], h[1], h[0])):
        d = np.fft.fftfreq(n_) * n_ * hi          # signed grid offsets in metres
        axes.append(d)
    dz, dy, dx = np.meshgrid(*axes, indexing="ij")
    r2 = dx * dx + dy * dy + dz * dz
    ker = np.exp(-2.0 * r2 / (corr_len * corr_len))
    ker /= math.sqrt(float(np.sum(ker * ker)))
    # real-to-complex transforms on all host cores (scipy.fft)
    g = sfft.irfftn(sfft.rfftn(w, workers=-1) * sfft.rfftn(ker, workers=-1), s=w.shape, workers=-1)
    g = g[pad[2]:pad[2] + nz, pad[1]:pad[1] + ny, pad[0]:pad[0] + nx]
    mu = math.log(mean) - 0.5 * log_std * log_std
    return np.exp(mu + log_std * g).reshape(-1)


@dataclass
class Mesh:
    """Brick mesh plus nodal loads/supports in the reference's input order."""
    nx: int
    ny: int
    nz: int
    coords: np.ndarray            # (n_nodes, 3) float64
    enodes: np.ndarray            # (n_elems, 8) int64, 0-based node positions
    E: np.ndarray                 # (n_elems,) float64
    nu: np.ndarray                # (n_elems,) float64
    nl_node: np.ndarray           # (n_nl,) int64
    nl_type: np.ndarray           # (n_nl,) int64: 1 displacement, 3 force
    nl_dir: np.ndarray            # (n_nl,) int64: 1 ux, 2 uy, 3 uz
    nl_val: np.ndarray            # (n_nl,) float64
    meta: dict = field(default_factory=dict)

    @property
    def n_nodes(self) -> int:
        return int(self.coords.shape[0])

    @property
    def n_elems(self) -> int:
        return int(self.enodes.shape[0])

    @property
    def n_dof(self) -> int:
        return 3 * self.n_nodes

    def expected_nnz(self) -> int:
        """nnz of a structured brick mesh: 9 (3(nx+1)-2)(3(ny+1)-2)(3(nz+1)-2)/9
        node pairs -- i.e. 9 * prod(3 (n_a + 1) - 2)."""
        return 9 * (3 * (self.nx + 1) - 2) * (3 * (self.ny + 1) - 2) * (3 * (self.nz + 1) - 2)


def cantilever(nx: int, ny: int, nz: int, L: float = 2.0, B: float = 0.4, H: float = 0.4,
               E: float | np.ndarray = 30e9, nu: float = 0.2, total_load: float = -10000.0,
               distort: float = 0.0, seed: int = 0) -> Mesh:
    """Structured cantilever.  `distort` > 0 moves interior nodes by up to
    that fraction of the element size (seeded) to exercise non-box bricks."""
    xs = np.linspace(0.0, L, nx + 1)
    ys = np.linspace(0.0, B, ny + 1)
    zs = np.linspace(0.0, H, nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    coords = np.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], axis=1).astype(np.float64)
    if distort > 0.0:
        hx, hy, hz = L / nx, B / ny, H / nz
        ii = np.arange(coords.shape[0])
        i = ii % (nx + 1)
        j = (ii // (nx + 1)) % (ny + 1)
        k = ii // ((nx + 1) * (ny + 1))
        interior = (i > 0) & (i < nx) & (j > 0) & (j < ny) & (k > 0) & (k < nz)
        for a, hh in enumerate((hx, hy, hz)):
            d = counter_normal(seed, ii, stream=11 + a)
            coords[:, a] += np.where(interior, distort * hh * np.tanh(d), 0.0)
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ex = ex.transpose(2, 1, 0).reshape(-1)
    ey = ey.transpose(2, 1, 0).reshape(-1)
    ez = ez.transpose(2, 1, 0).reshape(-1)

    def nid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    enodes = np.stack([
        nid(ex, ey, ez), nid(ex + 1, ey, ez), nid(ex + 1, ey + 1, ez), nid(ex, ey + 1, ez),
        nid(ex, ey, ez + 1), nid(ex + 1, ey, ez + 1), nid(ex + 1, ey + 1, ez + 1), nid(ex, ey + 1, ez + 1),
    ], axis=1).astype(np.int64)
    ne = enodes.shape[0]
    Ev = np.full(ne, float(E)) if np.isscalar(E) else np.asarray(E, dtype=np.float64).reshape(ne)
    nuv = np.full(ne, float(nu))
    jj, kk = np.meshgrid(np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    jj = jj.T.reshape(-1)
    kk = kk.T.reshape(-1)
    fixed = nid(0, jj, kk).astype(np.int64)
    loaded = nid(nx, jj, kk).astype(np.int64)
    nf = fixed.size
    # supports first (node order, ux uy uz), then the loads
    nl_node = np.concatenate([np.repeat(fixed, 3), loaded])
    nl_type = np.concatenate([np.full(3 * nf, 1), np.full(loaded.size, 3)]).astype(np.int64)
    nl_dir = np.concatenate([np.tile([1, 2, 3], nf), np.full(loaded.size, 3)]).astype(np.int64)
    nl_val = np.concatenate([np.zeros(3 * nf), np.full(loaded.size, total_load / loaded.size)])
    return Mesh(nx, ny, nz, coords, enodes, Ev, nuv, nl_node.astype(np.int64), nl_type,
                nl_dir, nl_val.astype(np.float64),
                meta={"L": L, "B": B, "H": H, "total_load": total_load, "distort": distort})


def prescribed_cantilever(nx: int, ny: int, nz: int, E: float | np.ndarray = 30e9, tip_uz: float = -1e-4,
                          pull_ux: float = 2e-5, load: float = -2000.0) -> Mesh:
    """Cantilever driven by prescribed displacements: ux = pull_ux on the
    nodes of the plane x = L/nx, listed before the supports (so that it meets
    them in the reference's sequential order), the x = 0 supports, uz = tip_uz
    on the x = L face, and a force `load` (z) spread over the mid-length
    plane."""
    m = cantilever(nx, ny, nz, E=E)

    def plane(i):
        jj, kk = np.meshgrid(np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
        return (i + (nx + 1) * (jj.T.reshape(-1) + (ny + 1) * kk.T.reshape(-1))).astype(np.int64)

    p1, p0, pl, pm = plane(1), plane(0), plane(nx), plane(nx // 2)
    nl_node = np.concatenate([p1, np.repeat(p0, 3), pl, pm])
    nl_type = np.concatenate([np.full(p1.size, 1), np.full(3 * p0.size, 1), np.full(pl.size, 1),
                              np.full(pm.size, 3)]).astype(np.int64)
    nl_dir = np.concatenate([np.full(p1.size, 1), np.tile([1, 2, 3], p0.size), np.full(pl.size, 3),
                             np.full(pm.size, 3)]).astype(np.int64)
    nl_val = np.concatenate([np.full(p1.size, pull_ux), np.zeros(3 * p0.size), np.full(pl.size, tip_uz),
                             np.full(pm.size, load / pm.size)])
    return Mesh(nx, ny, nz, m.coords, m.enodes, m.E, m.nu, nl_node, nl_type, nl_dir, nl_val.astype(np.float64),
                meta=dict(m.meta, prescribed=True))


def config_mesh(name: str, seed: int = 2024) -> Mesh:
    """The benchmark's named workloads (2.0 x 0.4 x 0.4 m cantilever, nu 0.2)."""
    if name == "plain_cpu":          # 20x4x4, homogeneous
        return cantilever(20, 4, 4)
    if name in ("plain_gpu", "mixed"):
        nx, ny, nz = 200, 40, 40
        cov = 0.3 if name == "plain_gpu" else 0.6
        return cantilever(nx, ny, nz, E=lognormal_E(seed, nx * ny * nz, cov=cov))
    if name == "learned":
        nx, ny, nz = 128, 32, 32
        E = grf_E(seed, nx, ny, nz, (2.0 / nx, 0.4 / ny, 0.4 / nz))
        return cantilever(nx, ny, nz, E=E)
    raise KeyError(name)


BATCH_SIZES = ((32, 8, 8), (48, 12, 12), (64, 16, 16), (96, 24, 24))


def batch_meshes(n: int = 256, seed: int = 7, sizes=BATCH_SIZES, cov: float = 0.3):
    """Monte-Carlo batch: each sample's mesh drawn uniformly from `sizes`
    (same 2.0 x 0.4 x 0.4 m cantilever), its own i.i.d. lognormal E field."""
    pick = (counter_normal(seed, np.arange(n), stream=21) * 1e6).astype(np.int64) % len(sizes)
    out = []
    for i in range(n):
        nx, ny, nz = sizes[int(pick[i])]
        out.append(cantilever(nx, ny, nz, E=lognormal_E(seed * 1000003 + i, nx * ny * nz, cov=cov)))
    return out


def input_ids(mesh: Mesh, seed: int = 0):
    """Node, element and material-set IDs for an input file with IDs (the
    form femReadInput renumbers without -p, fem/fem_io.c:50-80): shuffled,
    gapped numbers, unique within each kind."""
    rng = np.random.default_rng(seed)
    nid = 7 + 10 * rng.permutation(mesh.n_nodes).astype(np.int64)
    eid = 3 + 4 * rng.permutation(mesh.n_elems).astype(np.int64)
    mid = 1000 + 3 * rng.permutation(mesh.n_elems).astype(np.int64)
    return nid, eid, mid


def write_fem_input(mesh: Mesh, path: str, material_per_element: bool = True, monte=None, ids=None) -> None:
    """Write `mesh` in the reference's preparsed input format (fem/feminput.5,
    read by femReadInput fem/fem_io.c:380-640 with -p): material type 1 with
    values {DENS, EX, NU, ALPHA, KXX, C, PRICE} (fem/fem_m001.c:39-41), one
    material set per element when moduli differ; numbers printed %.17g.
    `monte` = (input rows, output rows), each row (pos, type, item, subitem,
    subitemiter), appends the random-variable section fem_monte_read_data
    reads (fem/fem_mont.c:866-909).  `ids` = (node IDs, element IDs,
    material-set IDs) writes the form with IDs instead (no -p): element node
    lists, element material sets and load nodes then name IDs."""
    ne, nn = mesh.n_elems, mesh.n_nodes
    if ids is None:
        nid = np.arange(1, nn + 1)
        eid = np.arange(1, ne + 1)
        mid = np.arange(1, ne + 1)
        nref = np.arange(nn)           # what element and load lines name: positions
    else:
        nid, eid, mid = (np.asarray(a, dtype=np.int64) for a in ids)
        nref = nid
    lines = ["1 -1", "1 0 0", "0"]
    if material_per_element:
        lines.append(str(ne))
        lines += [f"{int(mid[e])} 1 7" for e in range(ne)]
        lines.append(str(7 * ne))
        lines += [f"2500 {mesh.E[e]:.17g} {mesh.nu[e]:.17g} 0 0 0 0" for e in range(ne)]
        mat = np.arange(ne) if ids is None else mid
    else:
        lines += ["1", f"{int(mid[0])} 1 7", "7", f"2500 {mesh.E[0]:.17g} {mesh.nu[0]:.17g} 0 0 0 0"]
        mat = np.zeros(ne, dtype=np.int64) if ids is None else np.full(ne, mid[0])
    lines.append(str(nn))
    lines += [f"{int(nid[p])} {c[0]:.17g} {c[1]:.17g} {c[2]:.17g}" for p, c in enumerate(mesh.coords)]
    lines.append(str(ne))
    lines += [f"{int(eid[e])} 9 0 {int(mat[e])} 8" for e in range(ne)]
    lines.append(str(8 * ne))
    lines += [" ".join(str(int(nref[v])) for v in row) for row in mesh.enodes]
    lines.append("0 0")
    lines.append(str(mesh.nl_node.size))
    lines += [f"{int(nref[n])} {int(t)} {int(d)} {v:.17g}" for n, t, d, v in
              zip(mesh.nl_node, mesh.nl_type, mesh.nl_dir, mesh.nl_val)]
    lines.append("0")        # element loads
    lines.append("0 0 0")    # time-dependent array data
    lines.append("0")        # dynamics
    if monte is not None:
        rows_in, rows_out = monte
        lines.append(f"{len(rows_in)} {len(rows_out)}")
        lines += [" ".join(str(int(v)) for v in row) for row in list(rows_in) + list(rows_out)]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
