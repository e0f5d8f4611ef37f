"""Models and samples of the Monte plugin golden cases (fem/fem_mont.c), shared
by make_golden.py (which runs the compiled reference on them) and the tests
(which run the GPU plugin and the oracle on them).

A case is a cantilever of bricks with one material set per element, written
in the preparsed input format with a random-variable section: rows
(pos, type, item, subitem, subitemiter) with the variable codes of
fem/fem_mont.h:30-50.
"""

from __future__ import annotations

import numpy as np

from ufem_amd.mesh import cantilever, lognormal_E, write_fem_input

# variable codes, fem/fem_mont.h:30-50
RS, MAT, N, NLD, EL, NLPOS, DAMP, MAX_D = range(8)
RES_D, RES_R, RES_E, RES_SUM_E, RES_MAX_E, RES_MIN_E, RES_FAIL_E = range(8, 15)
RES_D_MAX, RES_D_MIN, RES_D_SUM, RES_A_MAX, RES_FAIL = range(15, 20)
# value and result type ids (fem/fem_mat.h:40-78, fem/fem_eres.h)
MAT_EX, MAT_NU = 2, 5
RES_SX, RES_SXY, RES_SYZ, RES_EXY, RES_EPX = 7, 10, 11, 16, 73

NX, NY, NZ = 12, 3, 3


def base_mesh():
    ne = NX * NY * NZ
    return cantilever(NX, NY, NZ, E=lognormal_E(21, ne, cov=0.25))


def node(i, j, k):
    return i + (NX + 1) * (j + (NY + 1) * k)


def elem(i, j, k):
    return i + NX * (j + NY * k)


def _rows(mesh):
    tip = node(NX, 1, 2)
    mid = node(6, 1, 1)
    tip_entries = np.nonzero((mesh.nl_type == 3))[0]
    load_pos = int(tip_entries[3])
    ins = [
        (MAT, elem(0, 0, 0), MAT_EX, 0),
        (MAT, elem(5, 1, 2), MAT_EX, 0),
        (MAT, elem(11, 2, 2), MAT_NU, 0),
        # force, uz at a tip node whose position is also a load position: the
        # reference names an NLD variable by nlNode[item] (fem/fem_mont.c:200)
        (NLD, node(NX, 0, 0), 3, 3),
        (N, mid, 1, 0),                     # x of an interior node
        (MAX_D, 0, 0, 0),
        (DAMP, 1, 0, 0),
        (NLPOS, load_pos, 3, 3),
        (MAT, elem(3, 0, 0), MAT_EX, 1),   # repeating value past the set: ignored
    ]
    outs = [
        (RES_D, tip, 3, 0),
        (RES_D, mid, 1, 0),
        (RES_D, tip, 5, 0),                 # a rotation: no DOF, 0
        (RES_D_MAX, tip, 3, 0),
        (RES_D_MIN, tip, 3, 0),
        (RES_D_SUM, tip, 3, 0),
        (RES_A_MAX, tip, 3, 0),
        (RES_R, node(0, 0, 0), 3, 0),
        (RES_R, node(0, 3, 3), 1, 0),
        (RES_E, elem(0, 0, 0), RES_SX, 1),
        (RES_E, elem(0, 1, 2), RES_EXY, 19),
        (RES_E, elem(7, 2, 1), RES_SYZ, 5),
        (RES_E, elem(0, 0, 0), RES_EPX, 1),
        (RES_E, elem(0, 0, 0), RES_SX, 0),
        (RES_SUM_E, 0, RES_SX, 0),
        (RES_MAX_E, 0, RES_SX, 0),
        (RES_MIN_E, 0, RES_SXY, 0),
        (RES_FAIL_E, 0, RES_SX, 0),
        (RES_FAIL_E, 0, RES_EPX, 0),
        (RES_FAIL, 0, 0, 0),
    ]
    return ins, outs


def case_rows(tag):
    """(input rows, output rows) with their pos fields.  "mc_a": inputs at
    0..ilen-1 and outputs at ilen..ilen+olen-1, so that output k lands in
    ofld[k] (the reference takes output k's slot from row k of the whole
    list, fem/fem_mont.c:521); "mc_b": both numbered from 0, so that later
    outputs overwrite earlier ones there."""
    mesh = base_mesh()
    ins, outs = _rows(mesh)
    if tag == "mc_a":
        ri = [(p,) + r for p, r in enumerate(ins)]
        ro = [(len(ins) + p,) + r for p, r in enumerate(outs)]
    elif tag == "mc_b":
        ins, outs = ins[:3], outs[:5]
        ri = [(p,) + r for p, r in enumerate(ins)]
        ro = [(p,) + r for p, r in enumerate(outs)]
    else:
        raise KeyError(tag)
    return mesh, ri, ro


def samples(tag, n=3, seed=5):
    """n samples of the inputs: moduli x U(0.7, 1.3), nu 0.15-0.3, loads and
    coordinates perturbed"""
    mesh, ri, _ = case_rows(tag)
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        v = []
        for (_, t, it, sb, ir) in ri:
            if t == MAT and sb == MAT_EX:
                v.append(float(mesh.E[it] * rng.uniform(0.7, 1.3)))
            elif t == MAT and sb == MAT_NU:
                v.append(float(rng.uniform(0.15, 0.3)))
            elif t in (NLD, NLPOS):
                v.append(float(-10000.0 / 16 * rng.uniform(0.5, 2.0)))
            elif t == N:
                v.append(float(mesh.coords[it, sb - 1] + rng.uniform(-0.03, 0.03)))
            elif t == MAX_D:
                v.append(float(rng.uniform(1e-5, 1e-3)))
            else:
                v.append(float(rng.uniform(0.0, 1.0)))
        out.append(v)
    return out


def write_case(tag, path):
    mesh, ri, ro = case_rows(tag)
    write_fem_input(mesh, path, monte=(ri, ro))
    return mesh, ri, ro
