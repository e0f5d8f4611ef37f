"""CPU restatement of uFEM's Monte plugin bookkeeping (fem/fem_mont.c) over the
oracle's linear-static solve (oracle/ufem_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the GPU plugin
(ufem_amd/csrc/ufem_monte.c), never as part of the product path.  Pinned by
tests/test_oracle_monte.py against the output fields the compiled reference
(oracle/_ref/ufem_ref_monte) wrote for tests/golden/monte_golden.json.

The model is given as a Mesh with one material set per element (set e holds
{DENS, EX, NU, ALPHA, KXX, C, PRICE} of element e, as ufem_amd.mesh.
write_fem_input writes it) and the variable rows (pos, type, item, subitem,
subitemiter).
"""

from __future__ import annotations

import dataclasses

import numpy as np

import ufem_oracle as uo

KNOWN_DOFS = 7
FEM_ZERO = 1e-17
MAT1_VAL = [1, 2, 5, 38, 37, 39, 36]        # fem/fem_m001.c:39
MAT1_VAL_RP = [1, 2, 5, 38, 37, 39]         # fem/fem_m001.c:41
E9_RES_RP = [7, 8, 9, 11, 12, 10, 13, 14, 15, 17, 18, 16, 73, 74, 75, 77, 78, 76, 37, 38]  # fem/fem_e009.c:881
RES_LEN = 14 * 27


class MonteOracle:
    def __init__(self, mesh, rows_in, rows_out):
        self.mesh = mesh
        self.ri = [tuple(int(v) for v in r) for r in rows_in]
        self.ro = [tuple(int(v) for v in r) for r in rows_out]
        self.rows = self.ri + self.ro
        ne = mesh.n_elems
        # material sets as the input file holds them (7 values per set)
        self.mp_val = np.zeros(7 * ne)
        self.mp_val[0::7] = 2500.0
        self.mp_val[1::7] = mesh.E
        self.mp_val[2::7] = mesh.nu
        self.coords = mesh.coords.copy()
        self.nl_val = mesh.nl_val.copy()
        sup = np.nonzero(mesh.nl_type == 1)[0]
        # fem_sol_res_alloc (fem/fem_sol.c:676-722): zero-valued supports at init
        self.res_entry = [k for k, i in enumerate(sup) if mesh.nl_val[i] == 0.0]
        self.racc = np.zeros(len(self.res_entry))
        self.max_disp = 0.0
        self.fail_id = -1
        for k, r in enumerate(self.ro):
            if r[1] == 19:
                self.fail_id = k

    def mat_index(self, mp, vtype, rep):
        """femGetRepValIndex (fem/fem_accs.c:297-381) for a 7-value type-1 set"""
        if mp < 0 or mp >= self.mesh.n_elems:
            return -1
        if rep <= 0:
            if vtype not in MAT1_VAL:
                return -1
            i_pos = MAT1_VAL.index(vtype)
        else:
            if vtype not in MAT1_VAL_RP:
                return 0
            i_pos = MAT1_VAL_RP.index(vtype) + 6 * (rep - 1) + 7
        return -1 if i_pos >= 7 else 7 * mp + i_pos

    def _solve(self):
        m = dataclasses.replace(self.mesh, E=self.mp_val[1::7].copy(), nu=self.mp_val[2::7].copy(),
                                coords=self.coords.copy(), nl_val=self.nl_val.copy())
        s = uo.OracleSystem(m)
        s.assemble(3)
        rv, _, _ = s.solve(3, FEM_ZERO / 10000.0, 3 * s.n, 0)
        if rv != 0:
            s.close()
            return rv, None, None, None
        u = s.u()
        s.reactions(3)
        R, _, _ = s.R()
        res = s.element_results(3)
        s.close()
        return 0, u, R, res

    def solution(self, ifld):
        """monte_solution (fem/fem_mont.c:694-855) with if_type 0: (rv, ofld)"""
        il, ol = len(self.ri), len(self.ro)
        pos = [r[0] for r in self.rows]
        ofld = np.zeros(max([p + 1 for p in pos] + [ol]))
        for k in range(ol):
            ofld[pos[k]] = 0.0
        m = self.mesh
        for (p, t, it, sb, ir) in self.ri:
            v = float(ifld[p])
            if t == 1:
                k = self.mat_index(it, sb, ir)
                if -1 < k < self.mp_val.size:
                    self.mp_val[k] = v
            elif t == 2 and 0 <= it < m.n_nodes and 1 <= sb <= 3:
                self.coords[it, sb - 1] = v
            elif t == 3:
                hit = np.nonzero((m.nl_node == it) & (m.nl_type == sb) & (m.nl_dir == ir))[0]
                if hit.size:
                    self.nl_val[hit[0]] = v
            elif t == 5 and 0 <= it < m.nl_type.size and m.nl_type[it] == sb and m.nl_dir[it] == ir:
                self.nl_val[it] = v
            elif t == 7:
                self.max_disp = abs(v)
        rv, u, R, res = self._solve()
        if rv != 0:
            return rv, ofld
        self.racc += R[self.res_entry]
        rnode = m.nl_node[m.nl_type == 1][self.res_entry]
        rdof = m.nl_dir[m.nl_type == 1][self.res_entry]
        nd = KNOWN_DOFS * m.n_nodes

        def dof_map(q):
            if q < 0 or q >= nd:
                return 0
            return 3 * (q // KNOWN_DOFS) + q % KNOWN_DOFS + 1 if q % KNOWN_DOFS < 3 else 0

        def u_at(d):
            return 0.0 if d < 1 or d > u.size else float(u[d - 1])

        def eres(e, vtype, rep):
            if rep <= 0 or e < 0 or e >= m.n_elems or vtype not in E9_RES_RP:
                return 0.0
            i_pos = E9_RES_RP.index(vtype) + 20 * (rep - 1)
            if i_pos >= RES_LEN:
                return 0.0
            gp, t = divmod(i_pos, 20)
            return float(res[e, gp, t]) if t < 12 else 0.0

        def all_eres(vtype):
            return [eres(e, vtype, k) for e in range(m.n_elems) for k in range(21)]

        for k in range(ol):
            p, t, it, sb, ir = self.ro[k]
            o = pos[k]
            q = it * KNOWN_DOFS + sb - 1
            inside = 0 <= q < nd
            if t == 8:
                ofld[o] = u_at(dof_map(q)) if inside else 0.0
            elif t == 15 and inside:
                ofld[o] = max(ofld[o], u_at(dof_map(q)))
            elif t == 16 and inside:
                ofld[o] = min(ofld[o], u_at(dof_map(q)))
            elif t == 17 and inside:
                ofld[o] += dof_map(dof_map(q))
            elif t == 9:
                hit = np.nonzero((rnode == it) & (rdof == sb))[0]
                if hit.size:
                    ofld[o] = self.racc[hit[0]]
            elif t == 10:
                ofld[o] = eres(it, sb, ir)
            elif t == 11:
                ofld[o] = sum(all_eres(sb))
            elif t == 12:
                ofld[o] = max([0.0] + all_eres(sb))
            elif t == 13:
                ofld[o] = min([0.0] + all_eres(sb))
            elif t == 14:
                ofld[o] = 1.0 if any(v > 0.0 for v in all_eres(sb)) else 0.0
            elif t == 19 and self.fail_id > -1:
                # fem_asse_max_disp_simple; fem_asse_fail_cond fails at once for
                # a type-1 material (no yield stress: 0 < |0| is false)
                if self.max_disp > 0.0 and np.abs(u).max() >= self.max_disp:
                    ofld[o] = 1.0
                if m.n_elems > 0:
                    ofld[o] = 1.0
        return 0, ofld
