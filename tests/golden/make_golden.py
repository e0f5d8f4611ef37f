"""Write the golden vectors of tests/golden/ufem_golden.npz from the compiled
uFEM reference (oracle/_ref/ufem_ref_harness, built by `make -C oracle ref`
from the reference sources in this container).

Run:  python tests/golden/make_golden.py           (vectors)
      python tests/golden/make_golden.py --res     (result files het4_cg.res,
                                                    het8_fbc.res, het4_np.res)
      python tests/golden/make_golden.py --monte   (Monte plugin, monte_golden.json;
                                                    needs `make -C oracle ref-monte`)
Each case generates a synthetic cantilever with ufem_amd.mesh, writes it in the
reference's preparsed input format, runs the reference harness (serial, like
`fem -p`), and keeps small arrays: K pattern digests and products, full K for
the tiny meshes, F, u, reactions and solver iteration counts.
"""

from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from ufem_amd.mesh import cantilever, input_ids, lognormal_E, prescribed_cantilever  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ufem_ref_harness")


def pattern_digest(indptr: np.ndarray, cols: np.ndarray) -> str:
    """sha256 over the sorted (row, col) int64 pairs, 0-based."""
    n = indptr.size - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
    pairs = np.stack([rows, cols.astype(np.int64)], axis=1)
    return hashlib.sha256(np.ascontiguousarray(pairs).tobytes()).hexdigest()


def rowslot_to_csr(rowlen, cols1, vals):
    n = rowlen.size
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(rowlen, out=indptr[1:])
    rows = np.repeat(np.arange(n, dtype=np.int64), rowlen)
    order = np.lexsort((cols1, rows))
    return indptr, (cols1[order] - 1).astype(np.int64), vals[order]


def probe_vector(n: int) -> np.ndarray:
    """Deterministic probe w_i = sin(0.37 i + 0.1) used to fingerprint K."""
    return np.sin(0.37 * np.arange(n) + 0.1)


def csr_matvec(indptr, cols, vals, x):
    rows = np.repeat(np.arange(indptr.size - 1), np.diff(indptr))
    return np.bincount(rows, weights=vals * x[cols], minlength=indptr.size - 1)


def harness_cmd(fin, td, solver="cg", eps=1e-21, fbc=False, ids=None):
    cmd = [HARNESS, "-i", fin, "-o", td, "-solver", solver, "-eps", repr(eps)]
    if fbc:
        cmd.append("-fbc")
    if ids is not None:
        cmd.append("-np")
    return cmd


def run_case(mesh, tag, solver="cg", eps=1e-21, maxit=None, ke=(), full_k=False, threads=0, fbc=False,
             ids=None):
    with tempfile.TemporaryDirectory() as td:
        from ufem_amd.mesh import write_fem_input
        fin = os.path.join(td, "in.fem")
        write_fem_input(mesh, fin, ids=ids)
        cmd = harness_cmd(fin, td, solver, eps, fbc, ids)
        if maxit is not None:
            cmd += ["-maxit", str(maxit)]
        if threads:
            cmd += ["-t", str(threads)]
        for e in ke:
            cmd += ["-ke", str(e)]
        out = subprocess.run(cmd, check=True, capture_output=True, text=True)
        info = json.loads(out.stdout)

        def rd(name, dt):
            return np.fromfile(os.path.join(td, name), dtype=dt)

        res = {}
        n = info["n"]
        res[f"{tag}/n"] = np.array(n)
        res[f"{tag}/rv"] = np.array(info["rv"])
        res[f"{tag}/spmv_calls"] = np.array(info["spmv_calls"])
        res[f"{tag}/normA"] = np.array(info["normA"])
        res[f"{tag}/u"] = rd("u.f64", np.float64)
        res[f"{tag}/R"] = rd("R.f64", np.float64)
        res[f"{tag}/Rnode"] = rd("Rnode.i64", np.int64)
        res[f"{tag}/Rdof"] = rd("Rdof.i64", np.int64)
        res[f"{tag}/F1"] = rd("F1.f64", np.float64)
        if fbc or ids is not None:
            res[f"{tag}/nDOFfld"] = rd("nDOFfld.i64", np.int64)
        if ids is not None:
            for name in ("eNodesL", "eMP", "eRS", "nlNode"):
                res[f"{tag}/{name}"] = rd(f"{name}.i64", np.int64)
        for e in ke:
            res[f"{tag}/Ke_{e}"] = rd(f"Ke_{e}.f64", np.float64).reshape(24, 24)
        w = probe_vector(n)
        for kname in ("K0", "K1"):
            rl = rd(f"{kname}_rowlen.i64", np.int64)
            c1 = rd(f"{kname}_cols.i64", np.int64)
            v = rd(f"{kname}_vals.f64", np.float64)
            ip, cs, vs = rowslot_to_csr(rl, c1, v)
            res[f"{tag}/{kname}_rowlen"] = rl
            res[f"{tag}/{kname}_sha"] = np.array(pattern_digest(ip, cs))
            res[f"{tag}/{kname}_Kw"] = csr_matvec(ip, cs, vs, w)
            res[f"{tag}/{kname}_diag"] = vs[_diag_pos(ip, cs)]
            res[f"{tag}/{kname}_fro"] = np.array(float(np.sqrt(np.sum(vs * vs))))
            # first rows in slot (insertion) order
            k = int(rl[:12].sum())
            res[f"{tag}/{kname}_head_cols"] = c1[:k]
            res[f"{tag}/{kname}_head_vals"] = v[:k]
            if full_k and kname == "K0":
                res[f"{tag}/{kname}_indptr"] = ip
                res[f"{tag}/{kname}_cols"] = cs
                res[f"{tag}/{kname}_vals"] = vs
        return res


def _diag_pos(ip, cs):
    n = ip.size - 1
    out = np.empty(n, dtype=np.int64)
    for i in range(n):
        j = ip[i] + np.searchsorted(cs[ip[i]:ip[i + 1]], i)
        out[i] = j
    return out


def cases():
    """(tag, mesh, kwargs) of every golden case."""
    c1 = cantilever(20, 4, 4)
    het8 = cantilever(8, 2, 2, E=lognormal_E(7, 32, cov=0.3))
    het4 = cantilever(4, 2, 2, E=lognormal_E(11, 16, cov=0.6))
    dist = cantilever(6, 3, 3, distort=0.25, seed=5, E=lognormal_E(3, 54, cov=0.3))
    cube = cantilever(1, 1, 1, L=1.0, B=1.0, H=1.0)
    pd = prescribed_cantilever(12, 3, 3, E=lognormal_E(13, 108, cov=0.3))   # windows of the column zeroing stay inside K
    return [
        ("c1_cg", c1, dict(solver="cg", ke=(0, 319))),
        ("c1_cg_eps12", c1, dict(solver="cg", eps=1e-12)),
        ("c1_bicgs", c1, dict(solver="bicgs")),
        # threaded: slot order and sums vary at rounding level from run to run
        # (femMatPutAdd under per-row locks); the file keeps its first run
        ("c1_cg_t4", c1, dict(solver="cg", threads=4)),
        ("het8_cg", het8, dict(solver="cg", full_k=True)),
        ("het8_bicgs", het8, dict(solver="bicgs")),
        ("het4_cg", het4, dict(solver="cg", full_k=True)),
        ("dist_cg", dist, dict(solver="cg", full_k=True, ke=(0, 20, 53))),
        ("cube_cg", cube, dict(solver="cg", ke=(0,), full_k=True)),
        ("pd_cg", pd, dict(solver="cg", eps=1e-12, full_k=True)),
        # -fbc: zero-valued supports left out of K (femFastBC)
        ("het8_fbc", het8, dict(solver="cg", full_k=True, fbc=True)),
        ("pd_fbc", pd, dict(solver="cg", eps=1e-12, full_k=True, fbc=True)),
        # input with IDs, renumbered by femRenumIndex (no -p)
        ("het4_np", het4, dict(solver="cg", full_k=True, ids=input_ids(het4, 3))),
    ]


def write_res_golden(tag="het4_cg"):
    """The reference's own result file (femWriteRes, fem/fem_io.c:648-734) for
    one golden case, kept as text next to the vectors."""
    mesh, kw = {t: (m, k) for t, m, k in cases()}[tag]
    with tempfile.TemporaryDirectory() as td:
        from ufem_amd.mesh import write_fem_input
        fin = os.path.join(td, "in.fem")
        write_fem_input(mesh, fin, ids=kw.get("ids"))
        cmd = harness_cmd(fin, td, kw.get("solver", "cg"), kw.get("eps", 1e-21), kw.get("fbc", False),
                          kw.get("ids"))
        subprocess.run(cmd, check=True, capture_output=True, text=True)
        with open(os.path.join(td, "result.res")) as f, open(os.path.join(HERE, f"{tag}.res"), "w") as g:
            g.write(f.read())
    print("wrote", os.path.join(HERE, f"{tag}.res"))


MONTE = os.path.join(ROOT, "oracle", "_ref", "ufem_ref_monte")


def write_monte_golden():
    """The reference's Monte plugin functions (fem/fem_mont.c with USE_MONTE,
    driven by oracle/ref_monte.c) on the cases of monte_cases.py: variable
    counts and names, and per sample the return code and the output field."""
    import monte_cases as mc
    if not os.path.exists(MONTE):
        raise SystemExit(f"missing {MONTE}: run `make -C oracle ref-monte` first")
    out = {}
    for tag in ("mc_a", "mc_b"):
        smp = mc.samples(tag)
        with tempfile.TemporaryDirectory() as td:
            fin = os.path.join(td, "in.fem")
            fs = os.path.join(td, "samples.txt")
            mc.write_case(tag, fin)
            with open(fs, "w") as f:
                f.write("\n".join(" ".join(repr(v) for v in row) for row in smp) + "\n")
            lines = subprocess.run([MONTE, fin, fs], check=True, capture_output=True,
                                   text=True).stdout.splitlines()
        itype, il, ol, ff = (int(v) for v in lines[0].split())
        rows = [ln.split() for ln in lines[1 + il + ol:]]
        out[tag] = {"interface": itype, "ilen": il, "olen": ol, "ffunc": ff,
                    "inames": lines[1:1 + il], "onames": lines[1 + il:1 + il + ol],
                    "samples": smp, "rv": [int(r[0]) for r in rows],
                    "ofld": [[float(v) for v in r[1:]] for r in rows]}
        print(tag, "samples", len(rows), "rv", out[tag]["rv"])
    path = os.path.join(HERE, "monte_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--res":
        for tag in ("het4_cg", "het8_fbc", "het4_np"):
            write_res_golden(tag)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--monte":
        write_monte_golden()
        return
    if not os.path.exists(HARNESS):
        raise SystemExit(f"missing {HARNESS}: run `make -C oracle ref` first")
    out = {}
    meta = {}
    for tag, mesh, kw in cases():
        res = run_case(mesh, tag, **kw)
        out.update(res)
        out[f"{tag}/E"] = mesh.E
        meta[tag] = {k: (list(v) if isinstance(v, tuple) else v) for k, v in kw.items() if k != "ids"}
        if "ids" in kw:
            meta[tag]["ids"] = [a.tolist() for a in kw["ids"]]
        print(tag, "n", int(res[f"{tag}/n"]), "spmv_calls", int(res[f"{tag}/spmv_calls"]),
              "rv", int(res[f"{tag}/rv"]))
    old_path = os.path.join(HERE, "ufem_golden.npz")
    if os.path.exists(old_path):   # keep the threaded case's first run (see cases())
        z = np.load(old_path, allow_pickle=False)
        for k in z.files:
            if k.startswith("c1_cg_t4/"):
                out[k] = z[k]
    out["_meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, "ufem_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "ufem_golden.npz"),
          os.path.getsize(os.path.join(HERE, "ufem_golden.npz")), "bytes")


if __name__ == "__main__":
    main()
