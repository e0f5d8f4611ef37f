"""ctypes binding of the oracle (oracle/libufem_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's CPU-baseline leg, always as the checker or the CPU baseline, never
as part of the product path.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libufem_oracle.so")
REF_HARNESS = os.path.join(_HERE, "_ref", "ufem_ref_harness")

_lib = None

P_D = C.POINTER(C.c_double)
P_L = C.POINTER(C.c_long)


class _Model(C.Structure):
    _fields_ = [("n_nodes", C.c_long), ("x", P_D), ("y", P_D), ("z", P_D),
                ("n_elems", C.c_long), ("enodes", P_L), ("E", P_D), ("nu", P_D),
                ("n_nl", C.c_long), ("nl_node", P_L), ("nl_type", P_L), ("nl_dir", P_L),
                ("nl_val", P_D)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle oracle`)")
        L = C.CDLL(LIB_PATH)
        L.uo_create.restype = C.c_void_p
        L.uo_create.argtypes = [C.POINTER(_Model), C.c_int]
        L.uo_free.argtypes = [C.c_void_p]
        L.uo_ndof.restype = C.c_long
        L.uo_ndof.argtypes = [C.c_void_p]
        L.uo_nreact.restype = C.c_long
        L.uo_nreact.argtypes = [C.c_void_p]
        L.uo_element_stiffness.argtypes = [P_D, C.c_double, C.c_double, C.c_int, P_D]
        L.uo_fill_K.argtypes = [C.c_void_p, C.c_int]
        L.uo_add_loads.argtypes = [C.c_void_p]
        L.uo_add_disps.argtypes = [C.c_void_p]
        L.uo_solve.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_long, C.c_int, P_L, P_D]
        L.uo_reactions.argtypes = [C.c_void_p, C.c_int]
        L.uo_element_results.argtypes = [C.c_void_p, C.c_int, P_D]
        L.uo_K_nnz.restype = C.c_long
        L.uo_K_nnz.argtypes = [C.c_void_p]
        L.uo_get_K.argtypes = [C.c_void_p, P_L, P_L, P_D]
        L.uo_get_F.argtypes = [C.c_void_p, P_D]
        L.uo_get_u.argtypes = [C.c_void_p, P_D]
        L.uo_set_u.argtypes = [C.c_void_p, P_D]
        L.uo_get_R.argtypes = [C.c_void_p, P_D, P_L, P_L]
        L.uo_get_dofmap.argtypes = [C.c_void_p, P_L]
        L.uo_mat_norm.restype = C.c_double
        L.uo_mat_norm.argtypes = [C.c_void_p]
        L.uo_spmv.argtypes = [C.c_void_p, P_D, P_D]
        L.uo_timings.argtypes = [C.c_void_p, P_D]
        _lib = L
    return _lib


def _pd(a):
    return a.ctypes.data_as(P_D)


def _pl(a):
    return a.ctypes.data_as(P_L)


def element_stiffness(xyz, E, nu, npts=3):
    """K_e (24x24) of one brick, xyz (8,3) in the reference node order."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(24)
    Ke = np.zeros(576)
    rv = lib().uo_element_stiffness(_pd(xyz), float(E), float(nu), int(npts), _pd(Ke))
    if rv != 0:
        raise ValueError(f"element stiffness failed ({rv})")
    return Ke.reshape(24, 24)


class OracleSystem:
    """One linear-static problem held by the oracle (row-slot K, F, u, R)."""

    def __init__(self, mesh, threads=1):
        self._keep = [np.ascontiguousarray(mesh.coords[:, i]) for i in range(3)]
        self._keep += [np.ascontiguousarray(mesh.enodes.reshape(-1), dtype=np.int64),
                       np.ascontiguousarray(mesh.E, dtype=np.float64),
                       np.ascontiguousarray(mesh.nu, dtype=np.float64),
                       np.ascontiguousarray(mesh.nl_node, dtype=np.int64),
                       np.ascontiguousarray(mesh.nl_type, dtype=np.int64),
                       np.ascontiguousarray(mesh.nl_dir, dtype=np.int64),
                       np.ascontiguousarray(mesh.nl_val, dtype=np.float64)]
        k = self._keep
        self._m = _Model(mesh.n_nodes, _pd(k[0]), _pd(k[1]), _pd(k[2]), mesh.n_elems, _pl(k[3]),
                         _pd(k[4]), _pd(k[5]), k[6].size, _pl(k[6]), _pl(k[7]), _pl(k[8]), _pd(k[9]))
        self.h = lib().uo_create(C.byref(self._m), int(threads))
        if not self.h:
            raise MemoryError("uo_create failed")
        self.n = lib().uo_ndof(self.h)
        self.nreact = lib().uo_nreact(self.h)

    def close(self):
        if self.h:
            lib().uo_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rv, what):
        if rv != 0:
            raise RuntimeError(f"oracle {what} failed ({rv})")

    def fill_K(self, npts=3):
        self._chk(lib().uo_fill_K(self.h, npts), "fill_K")

    def add_loads(self):
        self._chk(lib().uo_add_loads(self.h), "add_loads")

    def add_disps(self):
        self._chk(lib().uo_add_disps(self.h), "add_disps")

    def assemble(self, npts=3):
        self.fill_K(npts)
        self.add_loads()
        self.add_disps()

    def solve(self, solver=0, eps=1e-21, maxit=-1, rule=0):
        it = C.c_long(0)
        rr = C.c_double(0.0)
        rv = lib().uo_solve(self.h, solver, eps, maxit, rule, C.byref(it), C.byref(rr))
        return rv, it.value, rr.value

    def reactions(self, npts=3):
        self._chk(lib().uo_reactions(self.h, npts), "reactions")

    def element_results(self, npts=3):
        """[n_elems, npts^3, 12]: sigma x..xy, epsilon x..xy at each Gauss point"""
        out = np.zeros((self._m.n_elems, npts ** 3, 12))
        self._chk(lib().uo_element_results(self.h, npts, _pd(out)), "element results")
        return out

    def K_rowslot(self):
        nnz = lib().uo_K_nnz(self.h)
        rl = np.zeros(self.n, dtype=np.int64)
        cols = np.zeros(nnz, dtype=np.int64)
        vals = np.zeros(nnz)
        lib().uo_get_K(self.h, _pl(rl), _pl(cols), _pd(vals))
        return rl, cols, vals

    def K_csr(self):
        """Sorted CSR (0-based) of the row-slot matrix, explicit zeros kept."""
        return rowslot_to_csr(*self.K_rowslot())

    def F(self):
        out = np.zeros(self.n)
        lib().uo_get_F(self.h, _pd(out))
        return out

    def u(self):
        out = np.zeros(self.n)
        lib().uo_get_u(self.h, _pd(out))
        return out

    def set_u(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        lib().uo_set_u(self.h, _pd(u))

    def R(self):
        R = np.zeros(self.nreact)
        rn = np.zeros(self.nreact, dtype=np.int64)
        rd = np.zeros(self.nreact, dtype=np.int64)
        lib().uo_get_R(self.h, _pd(R), _pl(rn), _pl(rd))
        return R, rn, rd

    def mat_norm(self):
        return lib().uo_mat_norm(self.h)

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        lib().uo_spmv(self.h, _pd(x), _pd(y))
        return y

    def timings(self):
        t = np.zeros(5)
        lib().uo_timings(self.h, _pd(t))
        return dict(zip(("dofs", "fill", "bc", "solve", "react"), t.tolist()))


def rowslot_to_csr(rowlen, cols1, vals):
    """Row-slot (slot order, 1-based columns) -> sorted 0-based CSR."""
    n = rowlen.size
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(rowlen, out=indptr[1:])
    rows = np.repeat(np.arange(n, dtype=np.int64), rowlen)
    order = np.lexsort((cols1, rows))
    return indptr, (cols1[order] - 1).astype(np.int64), vals[order]
