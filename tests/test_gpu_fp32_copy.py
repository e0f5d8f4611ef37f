"""The mixed solvers' FP32 copy of K (`ufem_sys_spmv_f32`).

The copy rounds every value of K to FP32 (optionally keeping the remainder
K - fl32(K) rounded to bf16 beside it: 48-bit values) and adds to each
diagonal 3x3 block C_I = sym(sum_J (K_IJ - stored(K_IJ))), the symmetric
part of its block row's summed rounding errors.  Checked here:
- y = K32' x against a numpy model of that copy built from K's FP64 values;
- the property the correction is for: on a rigid translation, where the rows
  of a stiffness matrix cancel, the corrected product stays within a few
  percent of the plain FP32 product's error of the FP64 product (a correction
  with the wrong sign would double that error instead);
- symmetry of the corrected operator, (x, K32' y) = (K32' x, y);
- the mixed solvers still reach the FP64 solution, with the correction and
  without it (UFEM_NO_CORR32)."""

import numpy as np
import pytest

from ufem_amd import gpu as G
from ufem_amd.mesh import cantilever, lognormal_E

pytestmark = pytest.mark.gpu


def bf16_rne(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def fp32_copy_model(ip, cols, vals, n, rem=False):
    rows = np.repeat(np.arange(n), np.diff(ip))
    v32 = vals.astype(np.float32).astype(np.float64)
    if rem:
        v32 = v32 + bf16_rne(vals - v32)
    err = vals - v32
    C = np.zeros((n // 3, 3, 3))
    np.add.at(C, (rows // 3, rows % 3, cols % 3), err)
    C = 0.5 * (C + C.transpose(0, 2, 1))
    C = C.astype(np.float32).astype(np.float64)

    def apply(x):
        y = np.zeros(n)
        np.add.at(y, rows, v32 * x[cols])
        return y + np.einsum("iab,ib->ia", C, x.reshape(-1, 3)).ravel()
    return apply


@pytest.fixture(scope="module")
def system():
    mesh = cantilever(16, 4, 4, E=lognormal_E(21, 256, cov=0.3))
    g = G.GpuSystem(mesh)
    g.assemble(2)
    ip, cols, vals = g.K_csr()
    return g, ip, cols, vals


@pytest.mark.parametrize("rem", [False, True])
def test_fp32_copy_matches_its_model(system, rem):
    g, ip, cols, vals = system
    n = g.n
    model = fp32_copy_model(ip, cols, vals, n, rem)
    absK = np.abs(vals)
    rows = np.repeat(np.arange(n), np.diff(ip))
    rng = np.random.default_rng(7)
    for _ in range(3):
        x = rng.standard_normal(n)
        y = g.spmv_f32(x, rem)
        scale = np.zeros(n)
        np.add.at(scale, rows, absK * np.abs(x[cols]))
        err = np.abs(y - model(x)) / scale.max()
        print("fp32 copy vs model", err.max())
        assert err.max() <= 1e-14


def test_fp32_copy_keeps_translations(system, monkeypatch):
    """(the remainder form: tested through its model above and the solves
    below; its translation error is below FP64 rounding of the sums)"""
    g, ip, cols, vals = system
    n = g.n
    Kmax = np.abs(vals).max()
    ratios = []
    for k in range(3):
        t = np.zeros(n)
        t[k::3] = 1.0
        y64 = g.spmv(t)
        e_corr = np.abs(g.spmv_f32(t) - y64).max() / Kmax
        monkeypatch.setenv("UFEM_NO_CORR32", "1")
        e_plain = np.abs(g.spmv_f32(t) - y64).max() / Kmax
        monkeypatch.delenv("UFEM_NO_CORR32")
        print("translation", k, "corrected", e_corr, "plain fp32", e_plain)
        assert e_plain > 1e-9
        ratios.append(e_corr / e_plain)
    assert max(ratios) <= 0.1


def test_fp32_copy_is_symmetric(system):
    g, *_ = system
    rng = np.random.default_rng(8)
    x, y = rng.standard_normal(g.n), rng.standard_normal(g.n)
    a, b = x @ g.spmv_f32(y), g.spmv_f32(x) @ y
    assert abs(a - b) <= 1e-13 * (abs(a) + abs(b))


@pytest.mark.parametrize("solver,rem", [("mixed", False), ("mixed64", False), ("mixed64", True)])
@pytest.mark.parametrize("corr", [True, False])
def test_mixed_solves_with_and_without_correction(solver, rem, corr, monkeypatch):
    mesh = cantilever(24, 6, 6, E=lognormal_E(22, 864, cov=0.3))
    if not corr:
        monkeypatch.setenv("UFEM_NO_CORR32", "1")
    g = G.GpuSystem(mesh)
    g.assemble(2)
    ref = G.GpuSystem(mesh)
    ref.assemble(2)
    r = ref.solve("pcg", eps=1e-11, rule=1)
    info = g.solve(solver, eps=1e-10, rule=1, maxit=20000, f32_rem=rem)
    err = float(np.linalg.norm(g.u() - ref.u()) / np.linalg.norm(ref.u()))
    print(solver, "rem" if rem else "", "corr" if corr else "plain", "iters", info["iters"], "fp64", r["iters"],
          "u err", err)
    assert info["rv"] == 0 and info["true_relres"] <= 1e-10
    assert err <= 1e-8
