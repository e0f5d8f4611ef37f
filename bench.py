#!/usr/bin/env python
"""Benchmark of the linear-static solve path on one MI355X.

Workload (the metric's configuration): 2.0 x 0.4 x 0.4 m cantilever of
200 x 40 x 40 eight-node bricks (1,013,643 DOF), per-element Young's modulus
lognormal (mean 30 GPa, COV 0.3, seeded), nu = 0.2, x = 0 face fixed, -10 kN
in z spread over the x = L face.  K is assembled on the device (FP64), then
one step = one FP64 Jacobi-PCG solve from u = 0 to ||r||/||f|| <= 1e-8 with
K, f already resident in HBM.

Printed: one JSON line with the solve time (value, seconds), the roofline of
the solve's dominant kernel (PCG phase A: p' = z + beta p, q = K p', (p',q)),
SpMV bandwidth, the mixed-precision solve on the same system, and the CPU
baseline (the oracle's pthread restatement of uFEM's solver, bounded sample).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("solve s to ||r||/||f||<=1e-8: plain FP64 PCG vs learned+mixed PCG; SpMV GB/s")
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(mesh, gpu_iters, sample_iters=60):
    """uFEM's threaded CPU solver (oracle restatement, pthreads <= 8 like
    AF_MAX_THREADS) on the same system: K assembled with 27 Gauss points,
    `sample_iters` CG iterations timed, the solve time extrapolated to the
    iteration count the GPU needed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ufem_oracle as uo
    threads = max(1, min(8, os.cpu_count() or 1))
    t0 = time.perf_counter()
    s = uo.OracleSystem(mesh, threads=threads)
    s.fill_K(3)
    s.add_loads()
    s.add_disps()
    t_asm = time.perf_counter() - t0
    t1 = time.perf_counter()
    rv, it, rr = s.solve(0, 1e-8, sample_iters, 1)
    t_cg = time.perf_counter() - t1
    per_it = t_cg / max(it, 1)
    s.close()
    return {
        "value": per_it * gpu_iters,
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle pthread restatement of femEqsCGwJ/fem_fill_K, {threads} threads, on the full "
                   f"{mesh.n_dof}-DOF system: assembly (27 Gauss points) {t_asm:.2f} s, {it} CG iterations "
                   f"{t_cg:.2f} s ({per_it * 1e3:.1f} ms/iteration), solve extrapolated to the GPU's "
                   f"{gpu_iters} iterations"),
        "assembly_s": t_asm,
        "ms_per_iteration": per_it * 1e3,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--exec", default="streamed", choices=["streamed", "persistent", "fused", "classic"])
    ap.add_argument("--size", default="200,40,40", help="bricks nx,ny,nz (default: the metric's mesh)")
    ap.add_argument("--seed", type=int, default=2024)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards by slabs, see DESIGN.md, but this build runs on one GPU)", file=sys.stderr)
        sys.exit(2)

    from ufem_amd import gpu as G
    from ufem_amd.mesh import cantilever, lognormal_E

    nx, ny, nz = (int(v) for v in args.size.split(","))
    t0 = time.perf_counter()
    mesh = cantilever(nx, ny, nz, E=lognormal_E(args.seed, nx * ny * nz, cov=0.3))
    log(f"mesh {nx}x{ny}x{nz}: {mesh.n_dof} DOF in {time.perf_counter() - t0:.1f} s")
    dev = G.device_info()
    log("device", dev)

    t0 = time.perf_counter()
    g = G.GpuSystem(mesh)
    t_setup = time.perf_counter() - t0
    g.fill_K(2)
    fill_ms = g.timings()["fill_ms"]
    g.add_loads()
    g.add_disps()
    bc_ms = g.timings()["bc_ms"]
    log(f"setup {t_setup:.2f} s, assembly {fill_ms:.2f} ms ({mesh.n_elems / fill_ms * 1e3:.3e} elem/s), "
        f"supports {bc_ms:.3f} ms, {g.ncolors} colours")

    opts = dict(solver="pcg", eps=1e-8, rule=1, exec=args.exec)
    for _ in range(args.warmup):
        g.solve(**opts)
    walls, devs = [], []
    info = None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        info = g.solve(**opts)
        walls.append(time.perf_counter() - t0)
        devs.append(info["solve_ms"] / 1e3)
        if info["rv"] != 0:
            raise SystemExit(f"plain PCG did not converge: {info}")
    solve_s = float(np.mean(walls))
    log(f"plain FP64 PCG: {info['iters']} iterations, {solve_s:.4f} s wall, {np.mean(devs):.4f} s device, "
        f"true relres {info['true_relres']:.2e}")
    u64 = g.u()
    R, rn, rd = g.reactions()
    sum_Rz = float(R[rd == 3].sum())

    # dominant kernel of the solve, timed with HIP events on its own stream
    dom = {"streamed": ("cg1_w", "k_cg1_w<3> (w = K u with (w,u) partials)"),
           "classic": ("pcg_q", "k_pcg_q<3> (q = K p with (p,q) partials)"),
           "persistent": ("pcg_q", "k_pcg_q<3> (q = K p with (p,q) partials)"),
           "fused": ("pcg_a", "k_pcg_a<3> (p'=z+beta p, q=K p', (p',q))")}[args.exec]
    a_ms, a_bytes = g.time_kernel(dom[0], reps=50)
    s_ms, s_bytes = g.time_kernel("spmv", reps=50)
    nnz = g.nnz()
    csr_bytes = 12.0 * nnz + 4.0 * (g.n + 1) + 16.0 * g.n
    achieved = a_bytes / (a_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{dom[0]}_plain_gpu_traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "kernel": dom[1],
                "bytes_per_launch": a_bytes, "avg_launch_us": round(a_ms * 1e3, 2)}
    spmv = {"kernel": "k_spmv<3,double,double>", "avg_launch_us": round(s_ms * 1e3, 2),
            "bytes_per_launch": s_bytes, "GBps": round(s_bytes / (s_ms * 1e-3) / 1e9, 1),
            "csr_equivalent_bytes": csr_bytes,
            "csr_equivalent_GBps": round(csr_bytes / (s_ms * 1e-3) / 1e9, 1)}
    log("roofline", roofline)
    log("spmv", spmv)

    mixed = None
    if not args.no_mixed:
        g.solve(solver="mixed", eps=1e-8, rule=1, maxit=20000)      # warm-up
        t0 = time.perf_counter()
        mi = g.solve(solver="mixed", eps=1e-8, rule=1, maxit=20000)
        m_s = time.perf_counter() - t0
        um = g.u()
        mixed = {"solve_s": round(m_s, 4), "iters": mi["iters"], "outer": mi["outer"], "rv": mi["rv"],
                 "true_relres": mi["true_relres"],
                 "u_rel_err_vs_fp64": float(np.linalg.norm(um - u64) / np.linalg.norm(u64)),
                 "speedup_vs_fp64": round(solve_s / m_s, 3)}
        log("mixed", mixed)

    cpu = None
    if not args.no_cpu:
        log("CPU baseline sample ...")
        cpu = cpu_baseline(mesh, info["iters"])
        log("cpu", cpu["sample"])

    out = {
        "metric": METRIC,
        "value": round(solve_s, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(solve_s * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (f"cantilever 2.0x0.4x0.4 m, {nx}x{ny}x{nz} 8-node bricks, {mesh.n_dof} DOF, "
                                f"E lognormal mean 30 GPa COV 0.3 seed {args.seed}, nu 0.2; FP64 Jacobi-PCG "
                                f"to ||r||/||f|| <= 1e-8 from u=0, K assembled on device"),
                   "n_dof": mesh.n_dof, "nnz": nnz, "exec": args.exec},
        "iters": info["iters"],
        "true_relres": info["true_relres"],
        "dof_per_s": round(mesh.n_dof / solve_s, 1),
        "assembly": {"ms": round(fill_ms, 3), "elements_per_s": round(mesh.n_elems / fill_ms * 1e3, 1),
                     "supports_ms": round(bc_ms, 3)},
        "equilibrium_sum_Rz": sum_Rz,
        "roofline": roofline,
        "spmv": spmv,
        "mixed": mixed,
        "cpu_baseline": ({k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")} if cpu else None),
        "device": dev["name"],
    }
    print(json.dumps(out), flush=True)
    g.close()


if __name__ == "__main__":
    main()
