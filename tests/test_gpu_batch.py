"""Batched solve: every sample of one persistent launch matches its own
single-system solve (and the oracle), whatever the mesh mix."""

import os
import tempfile

import numpy as np
import pytest

import ufem_oracle as uo
from ufem_amd import gpu as G
from ufem_amd.mesh import batch_meshes, cantilever, lognormal_E

pytestmark = pytest.mark.gpu


def test_batch_matches_single_solves():
    meshes = batch_meshes(12, seed=3, sizes=((8, 2, 2), (12, 3, 3), (16, 4, 4)))
    b = G.GpuBatch(meshes)
    assert b.nclasses <= 3
    b.assemble(2)
    res = b.solve(eps=1e-12)
    print("batch", res["rv"], res["iters"], res["wg"], res["grid"])
    assert res["rv"] == 0 and np.all(res["status"] == 1)
    for i, m in enumerate(meshes):
        g = G.GpuSystem(m)
        g.assemble(2)
        info = g.solve("pcg", eps=1e-12, rule=1)
        ub, us = b.u(i), g.u()
        assert np.linalg.norm(ub - us) <= 1e-9 * np.linalg.norm(us), i
        assert abs(int(res["iters"][i]) - info["iters"]) <= 10
    # one sample against the oracle
    o = uo.OracleSystem(meshes[0])
    o.assemble(2)
    o.solve(0, 1e-14, 100000, 1)
    assert np.linalg.norm(b.u(0) - o.u()) <= 1e-9 * np.linalg.norm(o.u())


def test_batch_autotune_and_cache():
    meshes = batch_meshes(10, seed=5, sizes=((8, 2, 2), (16, 4, 4)))
    b = G.GpuBatch(meshes)
    b.assemble(2)
    with tempfile.TemporaryDirectory() as td:
        cache = os.path.join(td, "tune.txt")
        wg, grid = b.autotune(cache, budget_iters=20)
        assert wg in (256, 512, 1024) and grid > 0
        fields = open(cache).read().strip().split("\t")[-1].split()
        assert int(fields[0]) == wg and int(fields[1]) == grid and int(fields[2]) in (1, 2, 4, 8)
        b2 = G.GpuBatch(meshes)
        b2.assemble(2)
        assert b2.autotune(cache, budget_iters=20) == (wg, grid)   # from the cache
        res = b2.solve(eps=1e-10)
        assert res["rv"] == 0 and res["wg"] == wg


def test_batch_reports_unconverged():
    meshes = [cantilever(10, 2, 2, E=lognormal_E(1, 40)), cantilever(6, 2, 2)]
    b = G.GpuBatch(meshes)
    b.assemble(2)
    res = b.solve(eps=1e-12, maxit=3)
    assert res["rv"] == -1 and np.all(res["status"] == 3) and np.all(res["iters"] == 3)


@pytest.mark.parametrize("team", [2, 4, 8])
def test_batch_teams_agree(team):
    """Samples solved by teams of workgroups (team barrier, team partial sums)
    give the same solution as one workgroup per sample."""
    meshes = batch_meshes(9, seed=11, sizes=((8, 2, 2), (24, 6, 6)))
    b = G.GpuBatch(meshes)
    b.assemble(2)
    b.set_shape(512, 256, 1)
    r1 = b.solve(eps=1e-11)
    u1 = [b.u(i) for i in range(len(meshes))]
    b.set_shape(512, 256, team)
    r2 = b.solve(eps=1e-11)
    assert r1["rv"] == 0 and r2["rv"] == 0
    assert r2["team"].max() == team
    for i in range(len(meshes)):
        assert np.linalg.norm(b.u(i) - u1[i]) <= 1e-9 * np.linalg.norm(u1[i])
        assert abs(int(r1["iters"][i]) - int(r2["iters"][i])) <= 3


def test_stopped_batch_ends_normally():
    """The team waits of the batched kernel are bounded: a host that gives up
    waiting (timeout) and a wait limit too short for any team both end the
    launch normally, with the stopped samples reported (status 4 or not
    started) and AF_ERR; the same batch then solves in full again."""
    from ufem_amd.mesh import batch_meshes
    meshes = batch_meshes(64, seed=3)
    b = G.GpuBatch(meshes)
    b.assemble(2)
    b.set_shape(512, 64, 8)                 # teams of up to 8 workgroups on the largest meshes
    full = b.solve(eps=1e-8)
    assert full["rv"] == 0 and np.all(full["status"] == 1)
    b.set_limits(wait_ms=0, timeout_ms=1.0)   # the host stops the solve after 1 ms
    r = b.solve(eps=1e-8)
    print("host stop:", np.bincount(r["status"], minlength=5), G.last_error())
    assert r["rv"] == -1 and "host gave up" in G.last_error()
    assert np.any(r["status"] != 1)
    b.set_limits(wait_ms=1e-6, timeout_ms=0)  # every team wait gives up at once
    r = b.solve(eps=1e-8)
    print("wait limit:", np.bincount(r["status"], minlength=5), G.last_error())
    assert r["rv"] == -1 and np.any(r["status"] == 4)
    b.set_limits(0, 0)
    again = b.solve(eps=1e-8)
    assert again["rv"] == 0 and np.all(again["status"] == 1)
    for i in (0, 13, 63):
        g = G.GpuSystem(meshes[i])
        g.assemble(2)
        g.solve("pcg", eps=1e-12)
        assert np.linalg.norm(b.u(i) - g.u()) <= 1e-7 * np.linalg.norm(g.u())
        g.close()
    b.close()


@pytest.mark.parametrize("team,mode", [(1, 1), (4, 1), (1, 2), (4, 2)])
def test_batch_mixed_reaches_the_fp64_solution(team, mode):
    """Mixed modes (FP32 copies of K with their row-sum correction, 1: alone,
    2: with bf16 remainders; FP64 vectors and residual refreshes) converge on
    an FP64 residual to the FP64 batch's solution; switched off again, and
    after a new assembly, the batch solves exactly as before."""
    meshes = batch_meshes(9, seed=13, sizes=((8, 2, 2), (24, 6, 6)))
    b = G.GpuBatch(meshes)
    b.assemble(2)
    b.set_shape(512, 256, team)
    r64 = b.solve(eps=1e-10)
    u64 = [b.u(i) for i in range(len(meshes))]
    b.set_mixed(mode)
    r32 = b.solve(eps=1e-10)
    print("team", team, "mode", mode, "fp64 iters", r64["iters"], "mixed iters", r32["iters"])
    assert r64["rv"] == 0 and r32["rv"] == 0 and np.all(r32["status"] == 1)
    for i, m in enumerate(meshes):
        u = b.u(i)
        assert np.linalg.norm(u - u64[i]) <= 1e-8 * np.linalg.norm(u64[i]), i
        # (meshes this small converge superlinearly in FP64 -- 26 iterations
        # from 1e-6 to 1e-10 -- and the residual refreshes below the copy's
        # attainable accuracy, about 1e-5 for FP32 and 1e-8 for 48-bit
        # values, disturb that phase: a numpy model of the same iteration
        # takes 676 and 599 iterations for sample 0, the single-system mixed
        # solver 519 with the FP32 copy)
        assert r32["iters"][i] <= (3.5 if mode == 1 else 2.2) * r64["iters"][i] + 20, i
        assert r32["relres"][i] <= 1e-10
    # the true residual of one sample, from its own system
    g = G.GpuSystem(meshes[-1])
    g.assemble(2)
    f = g.spmv(u64[-1])
    rel = np.linalg.norm(g.spmv(b.u(len(meshes) - 1)) - f) / np.linalg.norm(f)
    assert rel <= 3e-10
    b.set_mixed(False)
    b.assemble(2)
    r = b.solve(eps=1e-10)
    assert np.array_equal(r["iters"], r64["iters"])
    for i in range(len(meshes)):
        assert np.array_equal(b.u(i), u64[i])
    b.set_mixed(mode)
    b.assemble(2)
    r = b.solve(eps=1e-10)
    assert np.array_equal(r["iters"], r32["iters"])
    with pytest.raises(ValueError):
        b.set_mixed(3)
