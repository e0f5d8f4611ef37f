"""The Monte plugin (fem/fem_mont.c) on CPU: the oracle's restatement of its
bookkeeping (oracle/monte_oracle.py) against the output fields the compiled
reference wrote (tests/golden/monte_golden.json, `make_golden.py --monte`),
and the library's input reader and variable names against the same file.

What the goldens pin: displacements at DOF-map positions (a rotation reads 0),
maxima and minima that start from 0, the displacement "sum" that adds DOF-map
entries, reactions summed over the samples, element results by
(type, repetition) in the brick's 20-value layout with repetition 0 and past
378 values reading 0, sums/extrema over every element and repetition 0..20,
the failure flags, material values by femGetRepValIndex (a repeating value past
the set ignored), and the output slots taken from the rows of the whole list
(case mc_b, where later outputs overwrite earlier ones)."""

import json
import os

import numpy as np
import pytest

import monte_oracle
from ufem_amd import gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "monte_golden.json")))

import sys  # noqa: E402
sys.path.insert(0, os.path.join(HERE, "golden"))
import monte_cases as mc  # noqa: E402


def compare(got, ref, rel=1e-10):
    got, ref = np.asarray(got, float), np.asarray(ref, float)
    scale = np.maximum(np.abs(ref), 1e-300)
    bad = np.abs(got - ref) > rel * scale
    # exact where the reference's value is an integer flag or a DOF-map sum
    ints = (ref == np.round(ref)) & (np.abs(ref) < 1e6) & (np.abs(ref) != 0) | (ref == 0)
    bad |= ints & (got != ref)
    return np.nonzero(bad)[0]


@pytest.mark.parametrize("tag", ["mc_a", "mc_b"])
def test_oracle_matches_reference_monte_outputs(tag):
    g = GOLD[tag]
    mesh, ri, ro = mc.case_rows(tag)
    o = monte_oracle.MonteOracle(mesh, ri, ro)
    assert o.fail_id == g["ffunc"]
    for k, ifld in enumerate(g["samples"]):
        rv, ofld = o.solution(ifld)
        assert rv == g["rv"][k]
        idx = compare(ofld[:g["olen"]], g["ofld"][k])
        assert idx.size == 0, (tag, k, idx, ofld[idx], np.asarray(g["ofld"][k])[idx])


def test_oracle_planted_slips_are_caught():
    """a reaction list cleared per sample, or element results read with the
    repetition index off by one, differ from the reference's fields"""
    g = GOLD["mc_a"]
    mesh, ri, ro = mc.case_rows("mc_a")
    o = monte_oracle.MonteOracle(mesh, ri, ro)
    o.solution(g["samples"][0])
    o.racc[:] = 0.0                       # as if cleared between samples
    _, ofld = o.solution(g["samples"][1])
    names = g["onames"]
    r = names.index("REAC_FX_196")
    assert abs(ofld[r] - g["ofld"][1][r]) > 0.1 * abs(g["ofld"][1][r])
    o2 = monte_oracle.MonteOracle(mesh, ri, ro)
    base = o2.solution(g["samples"][0])[1]
    e = names.index("ERES_E68_T11_I5")
    row = ro[e]
    ro_shift = list(ro)
    ro_shift[e] = row[:4] + (row[4] - 1,)
    o3 = monte_oracle.MonteOracle(mesh, ri, ro_shift)
    shifted = o3.solution(g["samples"][0])[1]
    assert abs(shifted[e] - g["ofld"][0][e]) > 1e-3 * abs(g["ofld"][0][e])
    assert compare(base[:g["olen"]], g["ofld"][0]).size == 0


@pytest.mark.parametrize("tag", ["mc_a", "mc_b"])
def test_library_reads_model_and_names_variables(tag, tmp_path):
    g = GOLD[tag]
    path = str(tmp_path / f"{tag}.fem")
    mc.write_case(tag, path)
    m = gpu.Monte(path)
    try:
        assert gpu.Monte.interface_type() == g["interface"] == 2
        assert (m.ilen, m.olen, m.ffunc) == (g["ilen"], g["olen"], g["ffunc"])
        assert m.inames() == g["inames"]
        assert m.onames() == g["onames"]
        assert gpu.lib().monte_ivar_name(m.param, m.ilen).decode() == "empty"
        assert gpu.lib().monte_ovar_name(m.param, -1).decode() == "empty"
    finally:
        m.close()


def test_library_rejects_inputs_off_the_path(tmp_path):
    """unreadable files, other element or load types, gravity"""
    mesh, ri, ro = mc.case_rows("mc_a")
    assert gpu.lib().monte_init_lib_stuff(str(tmp_path / "missing.fem").encode()) == -1
    assert gpu.lib().monte_init_lib_stuff(b"") == -1
    path = tmp_path / "m.fem"
    mc.write_case("mc_a", str(path))
    text = path.read_text().splitlines()
    # element type 9 -> 10 on the first element line (after nodes)
    nn = mesh.n_nodes
    first_elem = next(i for i, ln in enumerate(text) if ln.strip() == str(nn)) + nn + 2
    bad = list(text)
    bad[first_elem] = bad[first_elem].replace(" 9 ", " 10 ", 1)
    (tmp_path / "bad_elem.fem").write_text("\n".join(bad) + "\n")
    assert gpu.lib().monte_init_lib_stuff(str(tmp_path / "bad_elem.fem").encode()) == -1
    assert "type 9" in gpu.last_error()
    # gravity
    gi = text.index("0 0")
    bad = list(text)
    bad[gi] = "3 -9.81"
    (tmp_path / "grav.fem").write_text("\n".join(bad) + "\n")
    assert gpu.lib().monte_init_lib_stuff(str(tmp_path / "grav.fem").encode()) == -1
    assert "gravity" in gpu.last_error()
    # a truncated file
    (tmp_path / "short.fem").write_text("\n".join(text[:40]) + "\n")
    assert gpu.lib().monte_init_lib_stuff(str(tmp_path / "short.fem").encode()) == -1
    gpu.lib().monte_clean_lib_stuff(b"x")


def test_solution_needs_init():
    gpu.lib().monte_clean_lib_stuff(b"x")
    x = np.zeros(4)
    y = np.zeros(4)
    assert gpu.lib().monte_solution(b"x", gpu._pd(x), gpu._pd(y), 0) == -1
