#!/usr/bin/env python
"""Benchmark of the linear-static solve path on one MI355X.

Workload (the metric's configuration): 2.0 x 0.4 x 0.4 m cantilever of
200 x 40 x 40 eight-node bricks (1,013,643 DOF), per-element Young's modulus
lognormal (mean 30 GPa, COV 0.3, seeded), nu = 0.2, x = 0 face fixed, -10 kN
in z spread over the x = L face.  K is assembled on the device (FP64), then
one step = one FP64 Jacobi-PCG solve from u = 0 to ||r||/||f|| <= 1e-8 with
K, f already resident in HBM.

Printed: one JSON line with the solve time (value, seconds), the roofline of
the solve's dominant kernel (w = K u with (w, u) of the single-reduction PCG),
SpMV bandwidth, the mixed-precision solve on the same system (and on a COV
0.6 field), CG with the SSOR preconditioner and PCG with the two-level
(coarse-grid) correction on the same system, learned + mixed precision on the
same system (a network trained on the GPU from fresh lognormal fields; plain
FP64 PCG vs flexible CG with the network, FP64 and with the FP32 matrix, each
without and with the coarse correction), the learned preconditioner on
the 128 x 32 x 32 random-field cantilever (training on the GPU, then plain
vs learned flexible CG on an unseen field), the batched
Monte-Carlo solve of 256 uneven meshes in one launch, the Monte plugin
interface (one sample per call, and samples as slots of one launch, beside
the reference's own plugin functions when they are built), and the CPU
baseline (the oracle's pthread restatement of uFEM's solver, bounded sample).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu] [--no-mixed]
                    [--no-learned] [--no-batch] [--kernels-only] [--net-kernels-only]
                    [--twolevel-only] [--plain-only]
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


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(nx, ny, nz, seed, t1_iters=20):
    """uFEM's threaded CPU path (the oracle's pthread restatement of fem_dofs,
    fem_fill_K, fem_add_disps, femEqsCGwJ and the reactions pass, threads
    capped at 8 like AF_MAX_THREADS) on the same 1M-DOF system, every phase
    timed: the full solve to ||r||/||f|| <= 1e-8 (no extrapolation) at
    T = min(8, cores), then `t1_iters` CG iterations at T = 1 (K from 2x2x2
    points there: the per-iteration cost does not depend on K's values), and
    the 20x4x4 cantilever under the reference's own criterion (eps 1e-21,
    maxIt n), timed with the compiled reference when it is built.  Runs in a
    child process after the GPU legs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ufem_oracle as uo
    from ufem_amd.mesh import cantilever, lognormal_E
    ncpu = os.cpu_count() or 1
    threads = max(1, min(8, ncpu))
    mesh = cantilever(nx, ny, nz, E=lognormal_E(seed, nx * ny * nz, cov=0.3))
    t0 = time.perf_counter()
    s = uo.OracleSystem(mesh, threads=threads)
    t_dofs = time.perf_counter() - t0
    s.fill_K(3)
    s.add_loads()
    s.add_disps()
    rv, it, rr = s.solve(0, 1e-8, 20 * s.n, 1)
    s.reactions(3)
    tm = s.timings()
    R, _, rdof = s.R()
    s.close()
    out = {"threads": threads, "nproc": ncpu, "cpu_model": cpu_model(),
           "phases_s": {"dofs_alloc": round(t_dofs, 3), "fill_K_27pt": round(tm["fill"], 3),
                        "supports": round(tm["bc"], 4), "solve": round(tm["solve"], 3),
                        "reactions_pass": round(tm["react"], 3)},
           "iters": it, "rv": rv, "relres": rr, "sum_Rz": float(R[rdof == 3].sum()) if R.size else None}
    # T = 1 per iteration
    s1 = uo.OracleSystem(mesh, threads=1)
    s1.fill_K(2)
    s1.add_loads()
    s1.add_disps()
    t1 = time.perf_counter()
    _, it1, _ = s1.solve(0, 1e-30, t1_iters, 1)
    per1 = (time.perf_counter() - t1) / max(it1, 1)
    s1.close()
    out["t1_ms_per_iteration"] = round(per1 * 1e3, 2)
    out["t8_ms_per_iteration"] = round(tm["solve"] / max(it, 1) * 1e3, 2)
    # the reference's own criterion on the CPU-runnable configuration
    m1 = cantilever(20, 4, 4)
    harness = os.path.join(ROOT, "oracle", "_ref", "ufem_ref_harness")
    if os.path.exists(harness):
        import subprocess
        import tempfile
        from ufem_amd.mesh import write_fem_input
        with tempfile.TemporaryDirectory() as td:
            fin = os.path.join(td, "c1.fem")
            write_fem_input(m1, fin)
            r = subprocess.run([harness, "-i", fin, "-o", td, "-nodump-k"], capture_output=True, text=True,
                               timeout=300)
            j = json.loads(r.stdout)
        out["config1_reference_rule"] = {"kind": "reference", "iters": j["spmv_calls"] - 1, "rv": j["rv"],
                                         "solve_s": j["t_solve"], "eps": j["eps"], "maxit": j["maxit"]}
    else:
        o1 = uo.OracleSystem(m1)
        o1.assemble(3)
        rv1, it1c, _ = o1.solve(0, 1e-21, o1.n, 0)
        out["config1_reference_rule"] = {"kind": "port", "iters": it1c, "rv": rv1,
                                         "solve_s": o1.timings()["solve"], "eps": 1e-21, "maxit": o1.n}
        o1.close()
    return out


def start_cpu_baseline(args):
    """the CPU baseline as a child process (no GPU use), read by finish_cpu_baseline"""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", "--size", args.size,
           "--seed", str(args.seed)]
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)


def finish_cpu_baseline(proc, mesh, timeout=900):
    try:
        out, err = proc.communicate(timeout=timeout)
    except Exception:
        proc.kill()
        proc.communicate()
        log("CPU baseline did not finish in time")
        return None
    if proc.returncode != 0:
        log("CPU baseline failed:", err[-2000:])
        return None
    c = json.loads(out.strip().splitlines()[-1])
    ph = c["phases_s"]
    c.update({
        "value": ph["solve"], "unit": "s", "cores": c["threads"], "kind": "port",
        "sample": (f"oracle pthread restatement of uFEM's path (fem_dofs, fem_fill_K 27-point, fem_add_disps, "
                   f"femEqsCGwJ, reactions pass), {c['threads']} threads of {c['nproc']} ({c['cpu_model']}), "
                   f"full {mesh.n_dof}-DOF system solved to ||r||/||f|| <= 1e-8 in {c['iters']} iterations, "
                   f"timed whole (no extrapolation); phases {ph}; T=1 {c['t1_ms_per_iteration']} ms per "
                   f"iteration, T={c['threads']} {c['t8_ms_per_iteration']} ms; run in a child process "
                   f"after the GPU legs"),
    })
    return c


def bench_mixed_high_contrast(nx, ny, nz, seed):
    """FP64 Jacobi PCG vs mixed precision on the high-contrast field (lognormal
    E, COV 0.6), all to ||r||/||f|| <= 1e-8, with the displacement error of
    each against the FP64 solve: the configuration's method (FP32 matrix and
    inner CG to 1e-4 inside FP64 iterative refinement: the inner CG restarts
    at every refinement), the reliable-update form (refresh every factor 0.1,
    the direction kept), and the FP32 matrix with FP64 vectors."""
    from ufem_amd import gpu as G
    from ufem_amd.mesh import cantilever, lognormal_E
    mesh = cantilever(nx, ny, nz, E=lognormal_E(seed, nx * ny * nz, cov=0.6))
    g = G.GpuSystem(mesh)
    g.assemble(2)
    g.solve(solver="pcg", eps=1e-8, rule=1)
    t0 = time.perf_counter()
    r64 = g.solve(solver="pcg", eps=1e-8, rule=1)
    t64 = time.perf_counter() - t0
    u64 = g.u()
    out = {"workload": f"cantilever {nx}x{ny}x{nz}, E lognormal COV 0.6 seed {seed}",
           "fp64": {"iters": r64["iters"], "solve_s": round(t64, 4), "true_relres": r64["true_relres"]}}
    forms = {"refinement_inner_1e-4": dict(solver="mixed", inner_eps=1e-4, inner_restart=True),
             "reliable_update_0.1": dict(solver="mixed", inner_eps=0.1),
             "fp64_vectors_0.1": dict(solver="mixed64", inner_eps=0.1)}
    for name, kw in forms.items():
        g.solve(eps=1e-8, rule=1, maxit=60000, **kw)
        t0 = time.perf_counter()
        rm = g.solve(eps=1e-8, rule=1, maxit=60000, **kw)
        tm = time.perf_counter() - t0
        um = g.u()
        out[name] = {"iters": rm["iters"], "fp64_refreshes": rm["outer"], "solve_s": round(tm, 4), "rv": rm["rv"],
                     "true_relres": rm["true_relres"],
                     "u_rel_err_vs_fp64": float(np.linalg.norm(um - u64) / np.linalg.norm(u64)),
                     "speedup_vs_fp64": round(t64 / tm, 3)}
    g.close()
    return out


def khit_capacity(mesh, dmap, known=7):
    """row capacities as femKhit sizes them (fem/fem_mloc.c:83-116): 24 per
    element touching the row's node"""
    inc = np.bincount(np.asarray(mesh.enodes).ravel(), minlength=mesh.n_nodes)
    n = int(dmap.max())
    cap = np.zeros(n, dtype=np.int64)
    for d in range(3):
        r = dmap[np.arange(mesh.n_nodes) * known + d]
        sel = r > 0
        cap[r[sel] - 1] = 24 * inc[sel]
    return cap


def bench_dropin(mesh, native_s, native_ms, u64):
    """The drop-in boundary at the metric's size: fem_fill_K(AF_NO) on the GPU
    into the reference's row-slot K (DOF map with the zero supports left
    out, as -fbc), then femEqsCGwJ through the C-ABI on that row-slot matrix
    (relative rule, 1e-8): the first call builds the device copy (3x3 node
    blocks), a second re-reads the host matrix (default policy), a third
    uses the valid copy (policy 1).  Times: host wall of each call and the
    device time of the solve inside it, beside the native solve."""
    from ufem_amd import gpu as G
    dm = G.dof_map(mesh, fbc=True)
    n = int(dm.max())
    cap = khit_capacity(mesh, dm)
    t0 = time.perf_counter()
    K = G.RowSlotMatrix.empty(n, cap)
    t_alloc = time.perf_counter() - t0
    t0 = time.perf_counter()
    rv = G.fill_K(K, mesh, dm, 7, 2)
    t_fill = time.perf_counter() - t0
    if rv != 0:
        raise SystemExit(f"fill_K failed ({rv}): {G.last_error()}")
    fk = G.fill_K_timings()
    F = np.zeros(n)
    for nd, t, d, v in zip(mesh.nl_node, mesh.nl_type, mesh.nl_dir, mesh.nl_val):
        r = dm[int(nd) * 7 + int(d) - 1]
        if int(t) == 3 and r > 0:
            F[r - 1] += v
    b = G.Vec(F)
    G.set_stop_rule(1)
    calls = []
    for policy in (0, 0, 1):
        G.set_mirror_policy(policy)
        x = G.Vec(np.zeros(n))
        t0 = time.perf_counter()
        rv = G.femEqsCGwJ(K, b, x, 1e-8, n)
        wall = time.perf_counter() - t0
        info = G.last_solve()
        mi = K.mirror_info()
        calls.append({"policy": policy, "rv": rv, "wall_s": round(wall, 4), "solve_ms": round(info["solve_ms"], 2),
                      "iters": info["iters"], "build_ms": round(mi["build_ms"], 1),
                      "refresh_ms": round(mi["refresh_ms"], 1)})
    G.set_mirror_policy(0)
    G.set_stop_rule(0)
    full = np.zeros(mesh.n_nodes * 3)
    act = dm.reshape(mesh.n_nodes, 7)[:, :3].reshape(-1)
    full[act > 0] = x.data[act[act > 0] - 1]
    err = float(np.linalg.norm(full - u64) / np.linalg.norm(u64))
    K.release()
    return {"workload": "the metric's system as uFEM's row-slot K (zero supports left out, as -fbc)",
            "rows": n, "filled_slots": mi["nfill"], "block_size": mi["bs"],
            "rowslot_alloc_s": round(t_alloc, 3),
            "fill_K": {"wall_s": round(t_fill, 3), "setup_ms": round(fk["setup_ms"], 1),
                       "device_ms": round(fk["device_ms"], 2), "writeback_ms": round(fk["writeback_ms"], 1)},
            "calls": calls, "u_rel_err_vs_native": err,
            "native_solve_ms": round(native_ms, 2), "native_wall_s": round(native_s, 4),
            "solve_ms_ratio_to_native": round(calls[-1]["solve_ms"] / native_ms, 3)}


def bench_learned(fields, steps, seed, levels=3, fine=True, lr=3e-3, width=64):
    """Learned preconditioner on the 128 x 32 x 32 GRF cantilever (421,443 DOF):
    pairs from plain-PCG solves on `fields` fresh fields (32 each), `steps`
    Adam steps, then plain PCG and flexible CG with the network on a field
    not seen in training (seed `seed`), both to ||r||/||f|| <= 1e-8."""
    from ufem_amd import gpu as G
    from ufem_amd import learned as Lr
    from ufem_amd.mesh import cantilever
    nx, ny, nz = 128, 32, 32
    mesh = cantilever(nx, ny, nz)
    E_test = Lr.field_E(mesh, seed)
    mesh.E[:] = E_test
    s = G.GpuSystem(mesh)
    s.assemble(2)
    net = G.GpuNet(nx + 1, ny + 1, nz + 1, batch=4, levels=levels, fine=fine, width=width)
    net.init(1)
    net.reserve(32 * fields)
    pairs, collect_s = Lr.collect(s, net, mesh, range(1000, 1000 + fields), 32, every=8, log=log)
    loss, train_ms = Lr.train(net, steps, batch=4, lr=lr)
    log(f"learned: {pairs} pairs in {collect_s:.1f} s, {steps} steps in {train_ms / 1e3:.1f} s, "
        f"loss {loss[:20].mean():.3f} -> {loss[-50:].mean():.3f}")
    Lr.rebind(s, net, mesh, E_test)

    def timed(**kw):
        s.solve(**kw)
        t0 = time.perf_counter()
        r = s.solve(**kw)
        return r, time.perf_counter() - t0, s.u()

    rp, tp, up = timed(solver="pcg", eps=1e-8, rule=1)
    rl, tl, ul = timed(solver="learned", eps=1e-8, rule=1, net=net, fcg_m=1)
    stages = net.time_stages()
    s.set_coarse(nx + 1, ny + 1, nz + 1, 8, 8, 8)
    with_coarse = {}
    for name, kw in (("twolevel", dict(solver="twolevel")),
                     ("learned_coarse", dict(solver="learned", net=net, fcg_m=1, coarse=True))):
        r, t, u = timed(eps=1e-8, rule=1, **kw)
        with_coarse[name] = {"iters": r["iters"], "solve_s": round(t, 4), "rv": r["rv"],
                             "true_relres": r["true_relres"], "speedup_vs_plain": round(tp / t, 3),
                             "u_rel_err_vs_plain": float(np.linalg.norm(u - up) / np.linalg.norm(up))}
    out = {
        "workload": (f"cantilever 2.0x0.4x0.4 m, {nx}x{ny}x{nz} bricks, {mesh.n_dof} DOF, E = exp(GRF, "
                     f"corr 0.2 m, log-std 0.5, mean 30 GPa); test field seed {seed}, training fields "
                     f"1000..{1000 + fields - 1}"),
        "levels": levels, "node_grid_level": fine, "width": width, "parameters": int(net.P),
        "pairs": pairs, "collect_s": round(collect_s, 2), "train_steps": steps, "train_s": round(train_ms / 1e3, 2),
        "loss_first20": round(float(loss[:20].mean()), 4), "loss_last50": round(float(loss[-50:].mean()), 4),
        "plain": {"iters": rp["iters"], "solve_s": round(tp, 4), "true_relres": rp["true_relres"]},
        "learned": {"iters": rl["iters"], "solve_s": round(tl, 4), "fcg_m": 1, "fallback": rl["fallback"],
                    "true_relres": rl["true_relres"],
                    "u_rel_err_vs_plain": float(np.linalg.norm(ul - up) / np.linalg.norm(up))},
        "iteration_ratio": round(rp["iters"] / max(rl["iters"], 1), 3),
        "speedup_vs_plain": round(tp / tl, 3),
        "us_per_iteration": {"plain": round(tp / rp["iters"] * 1e6, 1), "learned": round(tl / rl["iters"] * 1e6, 1)},
        "network_stage_us": {k: round(v, 1) for k, v in stages.items()},
        "with_coarse_8x8x8": with_coarse,
    }
    net.close()
    s.close()
    return out


def bench_learned_mixed(g, mesh, fields, steps, seed_base=5000, cov=0.3, coarse=(8, 10, 10), levels=3, lr=3e-3,
                        fine=True, width=64):
    """Learned + mixed precision on the metric's own system `g` (the 1M-DOF
    lognormal cantilever, E = mesh.E): pairs from plain-PCG solves on `fields`
    fresh i.i.d. lognormal fields of the same distribution (seeds seed_base..),
    `steps` Adam steps, then plain FP64 PCG, flexible CG with the network
    (FP64) and flexible CG with the network and q = K p from the FP32 copy of
    K (FP64 residual refreshes), the same with the copy's bf16 remainders
    (48-bit values), all to ||r||/||f|| <= 1e-8 from u = 0, the network
    applied in FP64 (and, for comparison, the FP64 learned and the mixed solve
    with it on the bf16 matrix cores with f32 activations); then the same with
    the two-level correction on a coarse grid every `coarse` bricks (PCG
    without the network, FP32 matrix (and 48-bit) without the network, the
    network in FP64 and with the FP32 (and 48-bit) matrix)."""
    from ufem_amd import gpu as G
    from ufem_amd import learned as Lr
    from ufem_amd.mesh import lognormal_E
    E_test = mesh.E.copy()
    net = G.GpuNet(mesh.nx + 1, mesh.ny + 1, mesh.nz + 1, batch=4, levels=levels, fine=fine, width=width)
    net.init(1)
    net.reserve(32 * fields)
    pairs, collect_s = Lr.collect(g, net, mesh, range(seed_base, seed_base + fields), 32, every=8, log=log,
                                  field=lambda m, sd: lognormal_E(sd, m.n_elems, cov=cov))
    loss, train_ms = Lr.train(net, steps, batch=4, lr=lr)
    log(f"learned (1M DOF): {pairs} pairs in {collect_s:.1f} s, {steps} steps in {train_ms / 1e3:.1f} s, "
        f"loss {loss[:20].mean():.3f} -> {loss[-50:].mean():.3f}")
    Lr.rebind(g, net, mesh, E_test)

    def timed(**kw):
        g.solve(**kw)
        t0 = time.perf_counter()
        r = g.solve(**kw)
        return r, time.perf_counter() - t0, g.u()

    rp, tp, up = timed(solver="pcg", eps=1e-8, rule=1)
    rl, tl, ul = timed(solver="learned", eps=1e-8, rule=1, net=net, fcg_m=1)
    rm, tm, um = timed(solver="learned_mixed", eps=1e-8, rule=1, net=net, fcg_m=1)
    rr, tr, ur = timed(solver="learned_mixed", eps=1e-8, rule=1, net=net, fcg_m=1, f32_rem=True)
    # the same network applied on the bf16 matrix cores with f32 activations
    # (its rounding noise makes the iteration count depend on the load's scale)
    rf, tf, _ = timed(solver="learned", eps=1e-8, rule=1, net=net, fcg_m=1, net_prec="f32")
    rfm, tfm, _ = timed(solver="learned_mixed", eps=1e-8, rule=1, net=net, fcg_m=1, net_prec="f32")
    stages = net.time_stages64()
    stages32 = net.time_stages()

    def err(u):
        return float(np.linalg.norm(u - up) / np.linalg.norm(up))

    # the same solves with the two-level (coarse-grid) correction added; the
    # coarse matrix is assembled and inverted inside each timed solve
    nc = g.set_coarse(mesh.nx + 1, mesh.ny + 1, mesh.nz + 1, *coarse)
    variants = {
        "twolevel": dict(solver="twolevel"),
        "twolevel_mixed": dict(solver="twolevel_mixed"),
        "twolevel_mixed_rem": dict(solver="twolevel_mixed", f32_rem=True),
        "coarse_mixed": dict(solver="learned_mixed", coarse=True),
        "learned_coarse": dict(solver="learned", net=net, fcg_m=1, coarse=True),
        "learned_coarse_mixed": dict(solver="learned_mixed", net=net, fcg_m=1, coarse=True),
        "learned_coarse_f64": dict(solver="learned", net=net, fcg_m=1, coarse=True, net_prec="f64"),
        "learned_coarse_mixed_f64": dict(solver="learned_mixed", net=net, fcg_m=1, coarse=True, net_prec="f64"),
    }
    with_coarse = {"coarse_grid": {"factors": list(coarse), "coarse_dof": int(nc)}}
    for name, kw in variants.items():
        r, t, u = timed(eps=1e-8, rule=1, **kw)
        with_coarse[name] = {"iters": r["iters"], "solve_s": round(t, 4), "rv": r["rv"],
                             "true_relres": r["true_relres"], "u_rel_err_vs_fp64": err(u),
                             "speedup_vs_plain": round(tp / t, 3),
                             "us_per_iteration": round(t / max(r["iters"], 1) * 1e6, 1)}
        if "mixed" in name:
            with_coarse[name]["fp64_refreshes"] = r["outer"]
        log(name, with_coarse[name])

    out = {
        "workload": (f"the metric's system: {mesh.nx}x{mesh.ny}x{mesh.nz} bricks, {mesh.n_dof} DOF, E lognormal "
                     f"COV {cov}; training fields: {fields} fresh lognormal fields, seeds "
                     f"{seed_base}..{seed_base + fields - 1}"),
        "levels": levels, "node_grid_level": fine, "width": width, "parameters": int(net.P),
        "pairs": pairs, "collect_s": round(collect_s, 2), "train_steps": steps, "train_s": round(train_ms / 1e3, 2),
        "loss_first20": round(float(loss[:20].mean()), 4), "loss_last50": round(float(loss[-50:].mean()), 4),
        "plain": {"iters": rp["iters"], "solve_s": round(tp, 4), "true_relres": rp["true_relres"]},
        "learned": {"iters": rl["iters"], "solve_s": round(tl, 4), "fallback": rl["fallback"],
                    "true_relres": rl["true_relres"], "u_rel_err_vs_fp64": err(ul)},
        "learned_mixed": {"iters": rm["iters"], "solve_s": round(tm, 4), "fp64_refreshes": rm["outer"],
                          "fallback": rm["fallback"], "true_relres": rm["true_relres"], "u_rel_err_vs_fp64": err(um)},
        "learned_mixed_rem": {"matrix": "FP32 + bf16 remainder per value (48-bit), FP64 vectors",
                              "iters": rr["iters"], "solve_s": round(tr, 4), "fp64_refreshes": rr["outer"],
                              "fallback": rr["fallback"], "true_relres": rr["true_relres"],
                              "u_rel_err_vs_fp64": err(ur)},
        "iteration_ratio": round(rp["iters"] / max(rm["iters"], 1), 3),
        "fastest_learned_mixed": ({"form": "learned_mixed_rem", "speedup_vs_plain": round(tp / tr, 3)} if tr <= tm
                                  else {"form": "learned_mixed", "speedup_vs_plain": round(tp / tm, 3)}),
        "speedup_vs_plain": {"learned": round(tp / tl, 3), "learned_mixed": round(tp / tm, 3),
                             "learned_mixed_rem": round(tp / tr, 3)},
        "us_per_iteration": {"plain": round(tp / rp["iters"] * 1e6, 1), "learned": round(tl / rl["iters"] * 1e6, 1),
                             "learned_mixed": round(tm / rm["iters"] * 1e6, 1),
                             "learned_mixed_rem": round(tr / rr["iters"] * 1e6, 1)},
        "inference": "FP64 activations and sums on the FP64 matrix cores (the bf16 weights widened exactly)",
        "inference_f32": {"what": "the same network on the bf16 matrix cores with f32 activations",
                          "learned": {"iters": rf["iters"], "solve_s": round(tf, 4),
                                      "speedup_vs_plain": round(tp / tf, 3)},
                          "learned_mixed": {"iters": rfm["iters"], "solve_s": round(tfm, 4),
                                            "speedup_vs_plain": round(tp / tfm, 3)}},
        "network_stage_us": {k: round(v, 1) for k, v in stages.items()},
        "network_stage_us_f32": {k: round(v, 1) for k, v in stages32.items()},
        "with_coarse": with_coarse,
    }
    net.close()
    return out


def bench_batch(n_samples=256, seed=7, sequential_sample=16):
    """Batched Monte-Carlo: n_samples cantilevers drawn uniformly from four mesh
    sizes (32x8x8 .. 96x24x24), each with its own lognormal E field, solved in
    one persistent launch (autotuned shape) to 1e-8; compared with solving a
    sample of them one at a time with the streamed PCG."""
    from ufem_amd import gpu as G
    from ufem_amd.mesh import batch_meshes
    meshes = batch_meshes(n_samples, seed=seed)
    b = G.GpuBatch(meshes)
    asm_ms = b.assemble(2)
    t0 = time.perf_counter()
    wg, grid = b.autotune(None, 100)
    tune_s = time.perf_counter() - t0
    b.solve(eps=1e-8)
    r = b.solve(eps=1e-8)
    sm = r["sample_ms"]
    dofs = float(sum(m.n_dof for m in meshes))
    idx = np.linspace(0, n_samples - 1, sequential_sample).astype(int)
    seq_ms = []
    for i in idx:
        g = G.GpuSystem(meshes[i])
        g.assemble(2)
        g.solve(solver="pcg", eps=1e-8, rule=1)
        seq_ms.append(g.solve(solver="pcg", eps=1e-8, rule=1)["solve_ms"])
        g.close()
    seq_total = float(np.mean(seq_ms)) * n_samples
    # algorithmic bytes: per sample and iteration its SpMV (76 B per 3x3 block,
    # u read and w written) and the single-reduction update (7 vectors read,
    # 5 written) = 76 nnzb + 112 n
    nbytes = 0.0
    for m, it in zip(meshes, r["iters"]):
        nnzb = (3 * (m.nx + 1) - 2) * (3 * (m.ny + 1) - 2) * (3 * (m.nz + 1) - 2)
        nbytes += float(it) * (76.0 * nnzb + 112.0 * m.n_dof)
    gbps = nbytes / (r["total_ms"] * 1e-3) / 1e9
    out = {
        "samples": n_samples, "classes": b.nclasses, "assembly_ms": round(asm_ms, 2),
        "autotune": {"wg": r["wg"], "grid": r["grid"], "team_big": r["team_big"], "seconds": round(tune_s, 2)},
        "total_ms": round(r["total_ms"], 2), "samples_per_s": round(n_samples / (r["total_ms"] / 1e3), 2),
        "dof_per_s": round(dofs / (r["total_ms"] / 1e3), 1),
        "sample_ms_p50": round(float(np.percentile(sm, 50)), 2), "sample_ms_p99": round(float(np.percentile(sm, 99)), 2),
        "all_converged": bool(np.all(r["status"] == 1)), "iters_min_max": [int(r["iters"].min()), int(r["iters"].max())],
        "one_at_a_time_ms_estimate": round(seq_total, 1),
        "speedup_vs_one_at_a_time": round(seq_total / r["total_ms"], 3),
        "roofline": {"bound": "hbm", "algorithmic_bytes": nbytes, "achieved_GBps": round(gbps, 1),
                     "peak": HBM_PEAK_GBPS, "frac": round(gbps / HBM_PEAK_GBPS, 4),
                     "bytes_per_sample_iteration": "76 nnzb + 112 n"},
    }
    # mixed: FP32 copies of every sample's K (with the row-sum correction;
    # mixed48: plus bf16 remainders), FP64 vectors and residual refreshes,
    # each with its own autotuned shape
    check = idx[:4]
    u64 = [b.u(int(i)) for i in check]
    for name, mode in (("mixed", 1), ("mixed48", 2)):
        b.set_mixed(mode)
        b.autotune(None, 100)
        b.solve(eps=1e-8)
        rm = b.solve(eps=1e-8)
        uerr = max(float(np.linalg.norm(b.u(int(i)) - u) / np.linalg.norm(u)) for i, u in zip(check, u64))
        out[name] = {
            "total_ms": round(rm["total_ms"], 2), "samples_per_s": round(n_samples / (rm["total_ms"] / 1e3), 2),
            "vs_fp64_batch": round(r["total_ms"] / rm["total_ms"], 3),
            "all_converged": bool(np.all(rm["status"] == 1)),
            "iters_vs_fp64": round(float(rm["iters"].sum()) / float(r["iters"].sum()), 3),
            "u_rel_err_vs_fp64_max": uerr,
            "autotune": {"wg": rm["wg"], "grid": rm["grid"], "team_big": rm["team_big"]},
        }
    b.close()
    return out


def bench_monte(nx=40, ny=8, nz=8, n_samples=32, seed=11, sequential_sample=8, ref_sample=2):
    """The Monte plugin (monte_* of fem/fem_mont.c) on a 40x8x8 cantilever
    with one material set per element: 16 moduli and the tip load sampled,
    outputs tip uz, a root reaction, the largest and summed sigma_x.
    Samples/s one at a time (monte_solution: SSOR-CG, eps 1e-21, the
    reference's call) and as slots of one batched launch
    (monte_solution_batch); the reference's own plugin functions
    (oracle/_ref/ufem_ref_monte, when built) timed on a small sample."""
    import subprocess
    import tempfile
    from ufem_amd import gpu as G
    from ufem_amd.mesh import cantilever, lognormal_E, write_fem_input
    ne = nx * ny * nz
    mesh = cantilever(nx, ny, nz, E=lognormal_E(seed, ne, cov=0.3))
    rng = np.random.default_rng(seed)
    sets = rng.choice(ne, 16, replace=False)
    tip = nx + (nx + 1) * (ny // 2 + (ny + 1) * nz)
    ins = [(1, int(e), 2, 0) for e in sets] + [(3, tip, 3, 3)]
    outs = [(8, tip, 3, 0), (9, 0, 3, 0), (12, 0, 7, 0), (11, 0, 7, 0)]
    ri = [(p,) + r for p, r in enumerate(ins)]
    ro = [(len(ins) + p,) + r for p, r in enumerate(outs)]
    smp = np.concatenate([mesh.E[sets][None, :] * rng.uniform(0.6, 1.4, (n_samples, 16)),
                          rng.uniform(-1000.0, -200.0, (n_samples, 1))], axis=1)
    out = {"workload": f"cantilever {nx}x{ny}x{nz} bricks, {mesh.n_dof} DOF, {ne} material sets; "
                       f"16 moduli + tip load sampled; 4 outputs", "samples": n_samples}
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "monte.fem")
        write_fem_input(mesh, path, monte=(ri, ro))
        m = G.Monte(path)
        m.solution(smp[0])                         # device system built here
        t0 = time.perf_counter()
        for k in range(sequential_sample):
            rv, _ = m.solution(smp[k])
        seq_s = (time.perf_counter() - t0) / sequential_sample
        m.close()
        m = G.Monte(path)
        t0 = time.perf_counter()
        rv_b, fields, st = m.solution_batch(smp)
        batch_s = time.perf_counter() - t0
        m.close()
        out.update({"one_at_a_time_samples_per_s": round(1.0 / seq_s, 2), "rv": int(rv),
                    "batch_samples_per_s": round(n_samples / batch_s, 2), "batch_rv": int(rv_b),
                    "batch_vs_one_at_a_time": round(seq_s * n_samples / batch_s, 3)})
        ref = os.path.join(ROOT, "oracle", "_ref", "ufem_ref_monte")
        if os.path.exists(ref):
            fs = os.path.join(td, "samples.txt")
            with open(fs, "w") as f:
                f.write("\n".join(" ".join(repr(float(v)) for v in row) for row in smp[:ref_sample]) + "\n")
            t0 = time.perf_counter()
            res = subprocess.run([ref, path, fs], capture_output=True, text=True, timeout=600)
            ref_s = (time.perf_counter() - t0) / ref_sample
            if res.returncode == 0:
                rows = [ln.split() for ln in res.stdout.splitlines()[1 + len(ri) + len(ro):]]
                ref_tip = float(rows[0][1])
                out["reference_cpu"] = {"samples_per_s": round(1.0 / ref_s, 3), "cores": 1,
                                        "kind": "reference", "sample": f"{ref_sample} samples, serial",
                                        "tip_uz_rel_diff": abs(fields[0][0] - ref_tip) / abs(ref_tip)}
                out["speedup_one_at_a_time_vs_reference"] = round(ref_s / seq_s, 2)
                out["speedup_batch_vs_reference"] = round(ref_s * n_samples / batch_s, 2)
    return out


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
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--learn-fields", type=int, default=63)
    ap.add_argument("--learn-steps", type=int, default=10000)
    ap.add_argument("--learn-lr", type=float, default=1e-3)
    ap.add_argument("--net-node-grid", action="store_true",
                    help="the network with a node-grid level (default: the coarse-first form)")
    ap.add_argument("--net-levels", type=int, default=6, help="levels of the learned preconditioner's U-Net (2..8)")
    ap.add_argument("--net-width", type=int, default=16,
                    help="at most this many channels per layer of the U-Net (a multiple of 4, 4..64)")
    ap.add_argument("--learned-1m-only", action="store_true",
                    help="assemble the metric's system and run only the learned + mixed comparison")
    ap.add_argument("--kernels-only", action="store_true",
                    help="launch only the dominant solve kernel and a device copy of known size "
                         "(for hardware-counter runs)")
    ap.add_argument("--mixed64-only", action="store_true",
                    help="assemble the metric's system and run two solves with FP64 vectors over the FP32 copy "
                         "of K (for profiling)")
    ap.add_argument("--twolevel-only", action="store_true",
                    help="assemble the metric's system and run two two-level solves (for profiling)")
    ap.add_argument("--plain-only", action="store_true",
                    help="stop after the plain FP64 solves, the dominant-kernel timing and the results "
                         "pass (for kernel traces)")
    ap.add_argument("--net-kernels-only", action="store_true",
                    help="launch only the network's inference stages on the metric's node grid "
                         "(initial weights; for kernel traces)")
    ap.add_argument("--ssor-only", action="store_true",
                    help="two SSOR-CG solves of the metric's system (for kernel traces)")
    ap.add_argument("--learned-solve-only", action="store_true",
                    help="flexible-CG solves with the (untrained) network and with the two-level "
                         "correction on the metric's system (for kernel traces)")
    ap.add_argument("--cpu-baseline-only", action="store_true",
                    help="(internal) run the CPU baseline leg alone and print its JSON")
    args = ap.parse_args()
    if args.cpu_baseline_only:
        nx, ny, nz = (int(v) for v in args.size.split(","))
        print(json.dumps(cpu_baseline(nx, ny, nz, args.seed)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path would shard by slabs of the node grid, but it runs on one GPU)", file=sys.stderr)
        sys.exit(2)

    from ufem_amd import gpu as G
    from ufem_amd.mesh import cantilever, lognormal_E

    if args.net_kernels_only:
        # the network's inference stages alone on the metric's node grid, in
        # the bench's form (initial weights: the kernels do not depend on them)
        nx, ny, nz = (int(v) for v in args.size.split(","))
        mk = cantilever(nx, ny, nz, E=lognormal_E(args.seed, nx * ny * nz, cov=0.3))
        sk = G.GpuSystem(mk)
        sk.assemble(2)
        net = G.GpuNet(nx + 1, ny + 1, nz + 1, batch=1, levels=args.net_levels, fine=args.net_node_grid,
                       width=args.net_width)
        net.init(1)
        net.bind(sk, mk.E)
        log("network stages, bf16 matrix cores, f32 activations (us)",
            {k: round(v, 1) for k, v in net.time_stages(reps=20).items()})
        if not args.net_node_grid and args.net_width <= 32:
            log("network stages, FP64 (the solver's inference; us)",
                {k: round(v, 1) for k, v in net.time_stages64(reps=20).items()})
        net.close()
        sk.close()
        return

    nx, ny, nz = (int(v) for v in args.size.split(","))
    # the CPU baseline runs after the GPU legs: its threads beside them would
    # delay the host side of the short solves (two-level: about 20 ms)
    want_cpu = not args.no_cpu and not (args.kernels_only or args.twolevel_only or args.mixed64_only or args.plain_only
                                        or args.ssor_only
                                        or args.learned_solve_only or args.learned_1m_only)
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

    if args.kernels_only:
        g.time_kernel("copy", reps=5)
        g.time_kernel("cg1_w", reps=5)
        g.close()
        return
    if args.ssor_only:
        for _ in range(2):
            r = g.solve(solver="ssor", eps=1e-8, rule=1)
            log(f"SSOR-CG: {r['iters']} iterations, {r['solve_ms']:.2f} ms device")
        g.close()
        return
    if args.learned_solve_only:
        net = G.GpuNet(nx + 1, ny + 1, nz + 1, batch=1, levels=args.net_levels, fine=args.net_node_grid,
                       width=args.net_width)
        net.init(1)
        net.bind(g, mesh.E)
        g.set_coarse(nx + 1, ny + 1, nz + 1, 8, 10, 10)
        for name, kw in (("learned", {}), ("learned_mixed", {}), ("learned_mixed", {"f32_rem": True}),
                         ("learned_mixed", {"coarse": True})):
            r = g.solve(solver=name, eps=1e-8, rule=1, net=net, maxit=4000, **kw)
            log(f"{name} {kw}: {r['iters']} iterations, {r['solve_ms']:.2f} ms device")
        net.close()
        g.close()
        return
    if args.learned_1m_only:
        out = bench_learned_mixed(g, mesh, args.learn_fields, args.learn_steps, levels=args.net_levels,
                                  lr=args.learn_lr, fine=args.net_node_grid, width=args.net_width)
        print(json.dumps({"learned_mixed_1m": out}), flush=True)
        g.close()
        return
    if args.mixed64_only:
        # one warm and one profiled FP64-vector mixed solve of the metric's system
        for _ in range(2):
            r = g.solve(solver="mixed64", eps=1e-8, rule=1, maxit=20000)
            log(f"FP64-vector mixed PCG: {r['iters']} iterations, {r['outer']} FP64 refreshes, "
                f"{r['solve_ms']:.2f} ms device")
        g.close()
        return
    if args.twolevel_only:
        # one warm and one profiled two-level solve of the metric's system
        g.set_coarse(nx + 1, ny + 1, nz + 1, 8, 10, 10)
        for _ in range(2):
            r = g.solve(solver="twolevel", eps=1e-8, rule=1)
            log(f"two-level PCG: {r['iters']} iterations, {r['solve_ms']:.2f} ms device")
        g.close()
        return
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
    tpath = os.path.join(ROOT, "profiles", f"k_{dom[0]}_cantilever_{nx}x{ny}x{nz}_hbm_traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    # the read rate a plain stream of the same matrix bytes reaches (the faster
    # of ordinary and non-temporal loads)
    reads = [(g.time_kernel(m, reps=50), m) for m in ("read", "read_nt")]
    (r_ms, r_bytes), r_mode = min(reads, key=lambda t: t[0][0])
    read_gbps = r_bytes / (r_ms * 1e-3) / 1e9
    log("stream read GB/s", {m: round(b / (ms * 1e-3) / 1e9, 1) for (ms, b), m in reads})
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "kernel": dom[1],
                "bytes_per_launch": a_bytes, "avg_launch_us": round(a_ms * 1e3, 2),
                "stream_read": {"kernel": f"k_stream_read<{'true' if r_mode == 'read_nt' else 'false'}> "
                                          "(matrix values + columns, no gathers)",
                                "bytes": r_bytes, "avg_launch_us": round(r_ms * 1e3, 2),
                                "GBps": round(read_gbps, 1)},
                "frac_of_stream_read": round(achieved / read_gbps, 4)}
    spmv = {"kernel": "k_spmv<3,double,double>", "avg_launch_us": round(s_ms * 1e3, 2),
            "bytes_per_launch": s_bytes, "GBps": round(s_bytes / (s_ms * 1e-3) / 1e9, 1),
            "csr_equivalent_bytes": csr_bytes,
            "csr_equivalent_GBps": round(csr_bytes / (s_ms * 1e-3) / 1e9, 1)}
    log("roofline", roofline)
    log("spmv", spmv)

    # results pass (strains and stresses at the 27 Gauss points of every
    # element, fem_fill_K(AF_YES)) on the solved u, written to HBM
    r_ms = g.time_results(3, reps=10)
    r_bytes = mesh.n_elems * 27 * 12 * 8.0
    results = {"ms": round(r_ms, 3), "elements_per_s": round(mesh.n_elems / r_ms * 1e3, 1),
               "write_GBps": round(r_bytes / (r_ms * 1e-3) / 1e9, 1)}
    log("results pass", results)
    if args.plain_only:
        g.close()
        return

    log("drop-in boundary ...")
    dropin = bench_dropin(mesh, solve_s, float(np.mean(devs)) * 1e3, u64)
    log("dropin", dropin)

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
        # FP64 vectors over the corrected FP32 copy of K (and the 48-bit copy)
        for name, rem in (("fp64_vectors", False), ("fp64_vectors_48bit", True)):
            g.solve(solver="mixed64", eps=1e-8, rule=1, maxit=20000, f32_rem=rem)
            t0 = time.perf_counter()
            mi = g.solve(solver="mixed64", eps=1e-8, rule=1, maxit=20000, f32_rem=rem)
            m_s = time.perf_counter() - t0
            um = g.u()
            mixed[name] = {"solve_s": round(m_s, 4), "iters": mi["iters"], "outer": mi["outer"], "rv": mi["rv"],
                           "true_relres": mi["true_relres"],
                           "u_rel_err_vs_fp64": float(np.linalg.norm(um - u64) / np.linalg.norm(u64)),
                           "speedup_vs_fp64": round(solve_s / m_s, 3)}
        log("mixed", mixed)

    # CG with the SSOR preconditioner (femEqsCGwSSOR) on the same system; the
    # first call also builds the colour order
    g.solve(solver="ssor", eps=1e-8, rule=1)
    t0 = time.perf_counter()
    si = g.solve(solver="ssor", eps=1e-8, rule=1)
    s_s = time.perf_counter() - t0
    ssor = {"iters": si["iters"], "colours": si["outer"], "solve_s": round(s_s, 4), "rv": si["rv"],
            "true_relres": si["true_relres"],
            "u_rel_err_vs_jacobi": float(np.linalg.norm(g.u() - u64) / np.linalg.norm(u64)),
            "speedup_vs_jacobi": round(solve_s / s_s, 3)}
    log("ssor", ssor)

    # PCG with the two-level correction (coarse grid every 8 x 10 x 10 bricks; the
    # coarse matrix is assembled and inverted inside the timed solve)
    nc = g.set_coarse(nx + 1, ny + 1, nz + 1, 8, 10, 10)
    twolevel = {"coarse_factors": [8, 10, 10], "coarse_dof": int(nc)}
    for name in ("twolevel", "twolevel_mixed"):
        g.solve(solver=name, eps=1e-8, rule=1)
        t0 = time.perf_counter()
        ti = g.solve(solver=name, eps=1e-8, rule=1)
        t_s = time.perf_counter() - t0
        rec = {"iters": ti["iters"], "solve_s": round(t_s, 4), "rv": ti["rv"], "true_relres": ti["true_relres"],
               "u_rel_err_vs_jacobi": float(np.linalg.norm(g.u() - u64) / np.linalg.norm(u64)),
               "speedup_vs_jacobi": round(solve_s / t_s, 3)}
        if name == "twolevel":
            twolevel.update(rec)
        else:
            rec["fp64_refreshes"] = ti["outer"]
            twolevel["mixed"] = rec
    log("twolevel", twolevel)

    mixed_cov06 = None
    if not args.no_mixed:
        mixed_cov06 = bench_mixed_high_contrast(nx, ny, nz, args.seed + 1)
        log("mixed, COV 0.6", mixed_cov06)

    learned_mixed = None
    if not args.no_learned:
        log("learned + mixed precision at the metric's size ...")
        learned_mixed = bench_learned_mixed(g, mesh, args.learn_fields, args.learn_steps, levels=args.net_levels,
                                            lr=args.learn_lr, fine=args.net_node_grid, width=args.net_width)
        log("learned_mixed", learned_mixed)

    learned = None
    if not args.no_learned:
        log("learned preconditioner ...")
        learned = bench_learned(args.learn_fields, args.learn_steps, 9999, levels=args.net_levels,
                                fine=args.net_node_grid, lr=args.learn_lr, width=args.net_width)
        log("learned", learned)

    batch = None
    if not args.no_batch:
        log("batched Monte-Carlo ...")
        batch = bench_batch()
        log("batch", batch)

    monte = None
    if not args.no_batch:
        log("Monte plugin ...")
        monte = bench_monte()
        log("monte", monte)

    cpu = None
    if want_cpu:
        log("CPU baseline (child process) ...")
        cpu = finish_cpu_baseline(start_cpu_baseline(args), mesh)
        if cpu:
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
        "mixed_cov06": mixed_cov06,
        "ssor": ssor,
        "twolevel": twolevel,
        "results_pass": results,
        "learned_mixed_1m": learned_mixed,
        "learned": learned,
        "batch": batch,
        "monte_plugin": monte,
        "dropin": dropin,
        "cpu_baseline": ({k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")} if cpu else None),
        "cpu_detail": ({k: cpu[k] for k in ("phases_s", "iters", "rv", "relres", "sum_Rz", "t1_ms_per_iteration",
                                              "t8_ms_per_iteration", "nproc", "cpu_model",
                                              "config1_reference_rule")} if cpu else None),
        "gpu_vs_cpu_solve": (round(cpu["value"] / solve_s, 1) if cpu else None),
        "device": dev["name"],
    }
    print(json.dumps(out), flush=True)
    g.close()


if __name__ == "__main__":
    main()
