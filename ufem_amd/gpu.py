"""ctypes binding of libufem_gpu.so (the hand-written HIP path).

The library must be built (`python -m ufem_amd.build`); loading fails loudly
otherwise -- there is no CPU fallback on the product path.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libufem_gpu.so")

P_D = C.POINTER(C.c_double)
P_L = C.POINTER(C.c_long)
P_I = C.POINTER(C.c_int)
P_F = C.POINTER(C.c_float)

AF_OK, AF_ERR, AF_ERR_VAL, AF_ERR_TYP, AF_ERR_SIZ, AF_ERR_ZER = 0, -1, -3, -5, -9, -13
SOLVER = {"pcg": 0, "bicgstab": 1, "mixed": 2, "learned": 3, "learned_mixed": 4, "ssor": 5, "twolevel": 6,
          "twolevel_mixed": 7, "mixed64": 8}
EXEC = {"streamed": 0, "persistent": 1, "fused": 2, "classic": 3}


class tMatrix(C.Structure):
    """uFEM's matrix struct, fem/fem_math.h:47-57."""
    _fields_ = [("type", C.c_long), ("rows", C.c_long), ("cols", C.c_long), ("len", C.c_long),
                ("pos", P_L), ("data", P_D), ("frompos", P_L), ("defpos", P_L)]


class tVector(C.Structure):
    """uFEM's vector struct, fem/fem_math.h:60-67."""
    _fields_ = [("type", C.c_long), ("rows", C.c_long), ("len", C.c_long), ("pos", P_L),
                ("data", P_D)]


class _Mesh(C.Structure):
    _fields_ = [("n_nodes", C.c_long), ("x", P_D), ("y", P_D), ("z", P_D),
                ("n_elems", C.c_long), ("enodes", P_L), ("E", P_D), ("nu", P_D),
                ("n_nl", C.c_long), ("nl_node", P_L), ("nl_type", P_L), ("nl_dir", P_L),
                ("nl_val", P_D)]


class SolveOpts(C.Structure):
    _fields_ = [("solver", C.c_int), ("rule", C.c_int), ("eps", C.c_double), ("maxit", C.c_long),
                ("exec", C.c_int), ("chunk", C.c_int), ("inner_eps", C.c_double),
                ("inner_maxit", C.c_long), ("net", C.c_void_p), ("warm_start", C.c_int), ("fcg_m", C.c_int),
                ("coarse", C.c_int), ("inner_restart", C.c_int), ("f32_rem", C.c_int),
                ("net_prec", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_long), ("outer", C.c_long), ("fallback", C.c_int),
                ("relres", C.c_double), ("true_relres", C.c_double), ("solve_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run python -m ufem_amd.build)")
        L = C.CDLL(LIB_PATH)
        sig = {
            "ufem_sys_create": (C.c_void_p, [C.POINTER(_Mesh)]),
            "ufem_sys_destroy": (None, [C.c_void_p]),
            "ufem_sys_ndof": (C.c_long, [C.c_void_p]),
            "ufem_sys_nnz": (C.c_long, [C.c_void_p]),
            "ufem_sys_ncolors": (C.c_long, [C.c_void_p]),
            "ufem_sys_nreact": (C.c_long, [C.c_void_p]),
            "ufem_sys_fill_K": (C.c_int, [C.c_void_p, C.c_int]),
            "ufem_sys_set_E": (C.c_int, [C.c_void_p, P_D]),
            "ufem_net_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_int]),
            "ufem_net_destroy": (None, [C.c_void_p]),
            "ufem_net_nparams": (C.c_long, []),
            "ufem_net_create_levels": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "ufem_net_create_form": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "ufem_net_nparams_levels": (C.c_long, [C.c_int]),
            "ufem_net_nparams_form": (C.c_long, [C.c_int, C.c_int]),
            "ufem_net_create_width": (C.c_void_p, [C.c_int] * 7),
            "ufem_net_nparams_width": (C.c_long, [C.c_int, C.c_int, C.c_int]),
            "ufem_net_fine": (C.c_int, [C.c_void_p]),
            "ufem_net_levels": (C.c_int, [C.c_void_p]),
            "ufem_net_init": (C.c_int, [C.c_void_p, C.c_ulong]),
            "ufem_net_set_weights": (C.c_int, [C.c_void_p, P_F]),
            "ufem_net_get_weights": (C.c_int, [C.c_void_p, P_F]),
            "ufem_net_forward": (C.c_int, [C.c_void_p, C.c_int, P_F, P_F, P_F]),
            "ufem_net_backward": (C.c_int, [C.c_void_p, C.c_int, P_F, P_F, P_F, P_F]),
            "ufem_net_eval": (C.c_int, [C.c_void_p, C.c_long, C.c_long, P_D]),
            "ufem_net_get_pair": (C.c_int, [C.c_void_p, C.c_long, P_F, P_F, P_F, P_D, P_I]),
            "ufem_net_bind": (C.c_int, [C.c_void_p, C.c_void_p, P_D]),
            "ufem_net_reserve": (C.c_int, [C.c_void_p, C.c_long]),
            "ufem_net_npairs": (C.c_long, [C.c_void_p]),
            "ufem_net_collect": (C.c_long, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_void_p]),
            "ufem_net_train": (C.c_int, [C.c_void_p, C.c_long, C.c_int, C.c_double, C.c_ulong, P_D, P_D]),
            "ufem_net_time_stages": (C.c_int, [C.c_void_p, C.c_int, P_D]),
            "ufem_net_time_stages64": (C.c_int, [C.c_void_p, C.c_int, P_D]),
            "ufem_net_apply64": (C.c_int, [C.c_void_p, P_D, P_D]),
            "ufem_net_time_conv": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, P_I, P_I, C.c_int, C.c_int,
                                             C.c_int, P_D]),
            "ufem_net_probe_conv": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, P_I, P_I, C.c_int, C.c_int,
                                              P_F, P_F, P_F, P_F, P_F]),
            "ufem_sys_add_loads": (C.c_int, [C.c_void_p]),
            "ufem_sys_add_disps": (C.c_int, [C.c_void_p]),
            "ufem_sys_solve": (C.c_int, [C.c_void_p, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
            "ufem_sys_solve_noalloc": (C.c_int, [C.c_void_p, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
            "ufem_sys_reactions": (C.c_int, [C.c_void_p, P_D, P_L, P_L]),
            "ufem_sys_element_results": (C.c_int, [C.c_void_p, C.c_int, P_D]),
            "ufem_sys_write_results": (C.c_int, [C.c_void_p, C.c_char_p, C.c_long, C.c_long, C.c_double]),
            "ufem_sys_time_results": (C.c_double, [C.c_void_p, C.c_int, C.c_int]),
            "ufem_sys_get_u": (C.c_int, [C.c_void_p, P_D]),
            "ufem_sys_set_u": (C.c_int, [C.c_void_p, P_D]),
            "ufem_sys_get_F": (C.c_int, [C.c_void_p, P_D]),
            "ufem_sys_get_K": (C.c_int, [C.c_void_p, P_L, P_I, P_D]),
            "ufem_sys_spmv": (C.c_int, [C.c_void_p, P_D, P_D]),
            "ufem_sys_spmv_f32": (C.c_int, [C.c_void_p, P_D, P_D, C.c_int]),
            "ufem_sys_time_spmv": (C.c_int, [C.c_void_p, C.c_int, C.c_int, P_D, P_D]),
            "ufem_sys_timings": (None, [C.c_void_p, P_D]),
            "ufem_gpu_element_stiffness": (C.c_int, [C.c_long, P_D, P_D, P_D, C.c_int, P_D]),
            "ufem_gpu_set_stop_rule": (None, [C.c_int]),
            "ufem_gpu_last_solve": (None, [C.POINTER(SolveInfo)]),
            "ufem_gpu_last_error": (C.c_char_p, []),
            "ufem_gpu_device_info": (C.c_int, [C.c_char_p, C.c_int, P_I, P_L]),
            "ufem_batch_create": (C.c_void_p, [C.c_int, C.POINTER(_Mesh)]),
            "ufem_batch_destroy": (None, [C.c_void_p]),
            "ufem_batch_nclasses": (C.c_int, [C.c_void_p]),
            "ufem_batch_ndof": (C.c_long, [C.c_void_p, C.c_int]),
            "ufem_batch_assemble": (C.c_int, [C.c_void_p, C.c_int, P_D]),
            "ufem_batch_autotune": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, P_I, P_I]),
            "ufem_batch_shape": (None, [C.c_void_p, P_I, P_I, P_I]),
            "ufem_batch_set_shape": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
            "ufem_batch_teams": (C.c_int, [C.c_void_p, P_I]),
            "ufem_batch_solve": (C.c_int, [C.c_void_p, C.c_double, C.c_long, P_D, P_L, P_D, P_I, P_D]),
            "ufem_batch_get_u": (C.c_int, [C.c_void_p, C.c_int, P_D]),
            "ufem_batch_set_limits": (None, [C.c_void_p, C.c_double, C.c_double]),
            "ufem_batch_set_mixed": (C.c_int, [C.c_void_p, C.c_int]),
            "ufem_batch_abort": (None, [C.c_void_p]),
            "femEqsCGwJ": (C.c_int, [C.POINTER(tMatrix), C.POINTER(tVector), C.POINTER(tVector),
                                     C.c_double, C.c_long]),
            "femEqsBiCCSwJ": (C.c_int, [C.POINTER(tMatrix), C.POINTER(tVector), C.POINTER(tVector),
                                        C.c_double, C.c_long]),
            "femEqsCGwSSOR": (C.c_int, [C.POINTER(tMatrix), C.POINTER(tVector), C.POINTER(tVector),
                                        C.c_double, C.c_long]),
            "femMatVecMultBig": (C.c_int, [C.POINTER(tMatrix), C.POINTER(tVector), C.POINTER(tVector)]),
            "ufem_sys_update": (C.c_int, [C.c_void_p, C.POINTER(_Mesh)]),
            "ufem_sys_set_coarse": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "ufem_sys_coarse_size": (C.c_long, [C.c_void_p]),
            "ufem_sys_coarse_matrix": (C.c_int, [C.c_void_p, C.c_int, P_D]),
            "ufem_gpu_dense_spd_inverse": (C.c_int, [C.c_long, P_D, P_D]),
            "monte_dlib_interface_type": (C.c_long, []),
            "monte_nums_of_vars": (None, [C.c_char_p, P_L, P_L, P_L]),
            "monte_ivar_name": (C.c_char_p, [C.c_char_p, C.c_long]),
            "monte_ovar_name": (C.c_char_p, [C.c_char_p, C.c_long]),
            "monte_init_lib_stuff": (C.c_int, [C.c_char_p]),
            "monte_clean_lib_stuff": (C.c_int, [C.c_char_p]),
            "monte_solution": (C.c_int, [C.c_char_p, P_D, P_D, C.c_long]),
            "monte_solution_batch": (C.c_int, [C.c_char_p, C.c_long, P_D, P_D, C.c_long, P_I]),
            "femVecVecMultBig": (C.c_double, [C.POINTER(tVector), C.POINTER(tVector)]),
            "femVecNormBig": (C.c_double, [C.POINTER(tVector)]),
            "femMatNormBig": (C.c_double, [C.POINTER(tMatrix)]),
            "femVecLinCombBig": (C.c_int, [C.c_double, C.POINTER(tVector), C.c_double,
                                           C.POINTER(tVector), C.POINTER(tVector)]),
            "femMatSetZeroBig": (None, [C.POINTER(tMatrix)]),
            "femVecSetZeroBig": (None, [C.POINTER(tVector)]),
            "ufem_gpu_set_mirror_policy": (None, [C.c_int]),
            "ufem_gpu_mark_dirty": (None, [C.POINTER(tMatrix)]),
            "ufem_gpu_release": (None, [C.POINTER(tMatrix)]),
            "ufem_gpu_mirror_info": (C.c_int, [C.POINTER(tMatrix), P_D]),
            "ufem_gpu_fill_K": (C.c_int, [C.POINTER(tMatrix), C.POINTER(_Mesh), P_L, C.c_long, C.c_int]),
            "ufem_gpu_fill_K_timings": (None, [P_D]),
            "ufem_gpu_fill_K_release": (None, []),
            "ufem_model_read": (C.c_void_p, [C.c_char_p, C.c_int, P_I]),
            "ufem_model_free": (None, [C.c_void_p]),
            "ufem_model_mesh": (C.c_int, [C.c_void_p, C.POINTER(_Mesh)]),
            "ufem_model_node_ids": (P_L, [C.c_void_p]),
            "ufem_model_elem_ids": (P_L, [C.c_void_p]),
            "ufem_model_elem_mat": (P_L, [C.c_void_p]),
            "ufem_model_elem_rset": (P_L, [C.c_void_p]),
            "ufem_model_sol_id": (C.c_long, [C.c_void_p]),
            "ufem_model_sol_dep": (C.c_long, [C.c_void_p]),
            "ufem_sys_set_ids": (C.c_int, [C.c_void_p, P_L, P_L]),
            "ufem_sys_set_fbc": (C.c_int, [C.c_void_p, C.c_int]),
            "ufem_sys_dof_map": (C.c_long, [C.c_void_p, P_L]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().ufem_gpu_last_error().decode()


def device_info():
    name = C.create_string_buffer(256)
    ncu = C.c_int(0)
    hbm = C.c_long(0)
    rv = lib().ufem_gpu_device_info(name, 256, C.byref(ncu), C.byref(hbm))
    if rv != 0:
        raise RuntimeError(last_error())
    return {"name": name.value.decode(), "n_cu": ncu.value, "hbm_bytes": hbm.value}


def _pd(a):
    return a.ctypes.data_as(P_D)


def _pl(a):
    return a.ctypes.data_as(P_L)


def _pi(a):
    return a.ctypes.data_as(P_I)


class GpuSystem:
    """A linear-static problem resident on the GPU (K in SELL-64 3x3 blocks,
    F, u, supports).  Methods follow femSolve's steps."""

    def __init__(self, mesh):
        self._keep = [np.ascontiguousarray(mesh.coords[:, i], dtype=np.float64) for i in range(3)]
        self._keep += [np.ascontiguousarray(mesh.enodes.reshape(-1), dtype=np.int64),
                       np.ascontiguousarray(mesh.E, dtype=np.float64),
                       np.ascontiguousarray(mesh.nu, dtype=np.float64),
                       np.ascontiguousarray(mesh.nl_node, dtype=np.int64),
                       np.ascontiguousarray(mesh.nl_type, dtype=np.int64),
                       np.ascontiguousarray(mesh.nl_dir, dtype=np.int64),
                       np.ascontiguousarray(mesh.nl_val, dtype=np.float64)]
        k = self._keep
        m = _Mesh(mesh.n_nodes, _pd(k[0]), _pd(k[1]), _pd(k[2]), mesh.n_elems, _pl(k[3]), _pd(k[4]),
                  _pd(k[5]), k[6].size, _pl(k[6]), _pl(k[7]), _pl(k[8]), _pd(k[9]))
        self.h = lib().ufem_sys_create(C.byref(m))
        if not self.h:
            raise RuntimeError("ufem_sys_create: " + last_error())
        self.n = lib().ufem_sys_ndof(self.h)
        self.n_elems = mesh.n_elems
        self.nreact = lib().ufem_sys_nreact(self.h)
        self.ncolors = lib().ufem_sys_ncolors(self.h)

    def set_coarse(self, W, H, D, cx, cy, cz):
        """coarse node grid for the two-level correction (box meshes only)"""
        rv = lib().ufem_sys_set_coarse(self.h, W, H, D, cx, cy, cz)
        if rv != AF_OK:
            raise RuntimeError(f"ufem_sys_set_coarse = {rv}: {last_error()}")
        return lib().ufem_sys_coarse_size(self.h)

    def coarse_matrix(self, inverse=False):
        nc = lib().ufem_sys_coarse_size(self.h)
        A = np.zeros((nc, nc))
        self._chk(lib().ufem_sys_coarse_matrix(self.h, 1 if inverse else 0, _pd(A)), "coarse matrix")
        return A

    def update(self, mesh):
        """new coordinates, moduli and load/support values on the same mesh;
        F, K and the reactions start over (fill_K, add_loads, add_disps next)"""
        keep = []
        m = _mesh_struct(mesh, keep)
        rv = lib().ufem_sys_update(self.h, C.byref(m))
        if rv != AF_OK:
            raise RuntimeError(f"ufem_sys_update = {rv}: {last_error()}")

    def close(self):
        if getattr(self, "h", None):
            lib().ufem_sys_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rv, what):
        if rv != 0:
            raise RuntimeError(f"{what} failed ({rv}): {last_error()}")

    def fill_K(self, npts=2):
        self._chk(lib().ufem_sys_fill_K(self.h, npts), "fill_K")

    def set_E(self, E):
        """New per-element Young's moduli (then fill_K and add_disps again)."""
        self._keep_E = np.ascontiguousarray(E, dtype=np.float64)
        self._chk(lib().ufem_sys_set_E(self.h, _pd(self._keep_E)), "set_E")

    def add_loads(self):
        self._chk(lib().ufem_sys_add_loads(self.h), "add_loads")

    def add_disps(self):
        self._chk(lib().ufem_sys_add_disps(self.h), "add_disps")

    def assemble(self, npts=2):
        self.fill_K(npts)
        self.add_loads()
        self.add_disps()

    def solve(self, solver="pcg", eps=1e-8, rule=1, maxit=0, exec="streamed", chunk=0,
              inner_eps=0.0, inner_maxit=0, warm_start=False, net=None, fcg_m=0, coarse=False, check=True,
              inner_restart=False, f32_rem=False, net_prec=0):
        o = SolveOpts(SOLVER[solver], rule, eps, maxit, EXEC[exec], chunk, inner_eps, inner_maxit,
                      net.h if net is not None else None, 1 if warm_start else 0, fcg_m, 1 if coarse else 0,
                      1 if inner_restart else 0, 1 if f32_rem else 0, {0: 0, "auto": 0, "f64": 1, "f32": 2}[net_prec])
        info = SolveInfo()
        rv = lib().ufem_sys_solve(self.h, C.byref(o), C.byref(info))
        if check and rv not in (AF_OK, AF_ERR):
            raise RuntimeError(f"solve failed ({rv}): {last_error()}")
        d = info.as_dict()
        d["rv"] = rv
        return d

    def reactions(self):
        R = np.zeros(self.nreact)
        rn = np.zeros(self.nreact, dtype=np.int64)
        rd = np.zeros(self.nreact, dtype=np.int64)
        self._chk(lib().ufem_sys_reactions(self.h, _pd(R), _pl(rn), _pl(rd)), "reactions")
        return R, rn, rd

    def u(self):
        out = np.zeros(self.n)
        self._chk(lib().ufem_sys_get_u(self.h, _pd(out)), "get_u")
        return out

    def element_results(self, npts=3):
        """[n_elems, npts^3, 12]: sigma x..xy, epsilon x..xy at each Gauss point"""
        out = np.zeros((self.n_elems, npts ** 3, 12))
        self._chk(lib().ufem_sys_element_results(self.h, npts, _pd(out)), "element_results")
        return out

    def time_results(self, npts=3, reps=10):
        """average device ms of the results pass"""
        ms = lib().ufem_sys_time_results(self.h, npts, reps)
        if ms < 0:
            raise RuntimeError("time_results: " + last_error())
        return ms

    def set_ids(self, node_id, elem_id):
        """the input's node and element IDs for the result file"""
        self._ids = (np.ascontiguousarray(node_id, dtype=np.int64), np.ascontiguousarray(elem_id, dtype=np.int64))
        self._chk(lib().ufem_sys_set_ids(self.h, _pl(self._ids[0]), _pl(self._ids[1])), "set_ids")

    def set_fbc(self, fbc=True):
        """result file as the fem binary's -fbc writes it (zero supports left out)"""
        self._chk(lib().ufem_sys_set_fbc(self.h, 1 if fbc else 0), "set_fbc")

    def dof_map(self):
        """(nDOFfld with 7 entries per node, nDOFAct)"""
        m = np.zeros(7 * (self.n // 3), dtype=np.int64)
        act = lib().ufem_sys_dof_map(self.h, _pl(m))
        return m, int(act)

    def write_results(self, path, sol_id=1, dep_on=-1, sim_num=0.0):
        """the reference's result file (femWriteRes format) for the current u"""
        self._chk(lib().ufem_sys_write_results(self.h, path.encode(), sol_id, dep_on, sim_num), "write_results")

    def set_u(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._chk(lib().ufem_sys_set_u(self.h, _pd(u)), "set_u")

    def F(self):
        out = np.zeros(self.n)
        self._chk(lib().ufem_sys_get_F(self.h, _pd(out)), "get_F")
        return out

    def nnz(self):
        return lib().ufem_sys_nnz(self.h)

    def K_csr(self, values=True):
        nnz = self.nnz()
        ip = np.zeros(self.n + 1, dtype=np.int64)
        cols = np.zeros(nnz, dtype=np.int32)
        vals = np.zeros(nnz) if values else None
        self._chk(lib().ufem_sys_get_K(self.h, _pl(ip), _pi(cols), _pd(vals) if values else None), "get_K")
        return ip, cols.astype(np.int64), vals

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        self._chk(lib().ufem_sys_spmv(self.h, _pd(x), _pd(y)), "spmv")
        return y

    def spmv_f32(self, x, rem=False):
        """y = K x with the mixed solvers' FP32 copy of K (FP64 vectors; rem: with
        the bf16 remainders, 48-bit values)"""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        self._chk(lib().ufem_sys_spmv_f32(self.h, _pd(x), _pd(y), 1 if rem else 0), "spmv_f32")
        return y

    def time_kernel(self, mode="spmv", reps=20):
        """(avg ms, bytes) of one kernel: 'spmv', 'spmv32', 'pcg_a', 'pcg_q', 'cg1_w', 'copy', or 'read' /
        'read_nt' (a plain streaming read of the matrix's stored values and columns, with ordinary or
        non-temporal loads)"""
        m = {"spmv": 0, "spmv32": 1, "pcg_a": 2, "pcg_q": 3, "cg1_w": 4, "copy": 5, "read": 6, "read_nt": 7}[mode]
        ms = C.c_double(0)
        by = C.c_double(0)
        self._chk(lib().ufem_sys_time_spmv(self.h, reps, m, C.byref(ms), C.byref(by)), "time_kernel")
        return ms.value, by.value

    def timings(self):
        t = np.zeros(3)
        lib().ufem_sys_timings(self.h, _pd(t))
        return {"fill_ms": t[0], "bc_ms": t[1], "solve_ms": t[2]}


def _pf(a):
    return a.ctypes.data_as(P_F)


class GpuNet:
    """The learned preconditioner's U-Net (ufem_net_*) on a W x H x D node
    grid, with `levels` levels (2..8); fine=False: the coarse-first form
    (no node-grid level); width: at most this many channels per layer."""

    def __init__(self, W, H, D, batch=1, levels=3, fine=True, width=64):
        self.h = lib().ufem_net_create_width(W, H, D, batch, levels, 1 if fine else 0, width)
        if not self.h:
            raise RuntimeError("ufem_net_create: " + last_error())
        self.shape = (D, H, W)
        self.V = W * H * D
        self.levels = levels
        self.fine = fine
        self.width = width
        self.P = lib().ufem_net_nparams_width(levels, 1 if fine else 0, width)

    def close(self):
        if getattr(self, "h", None):
            lib().ufem_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rv, what):
        if rv != 0:
            raise RuntimeError(f"{what} failed ({rv}): {last_error()}")

    def init(self, seed=0):
        self._chk(lib().ufem_net_init(self.h, seed), "net init")

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.P
        self._chk(lib().ufem_net_set_weights(self.h, _pf(w)), "set_weights")

    def weights(self):
        w = np.zeros(self.P, dtype=np.float32)
        self._chk(lib().ufem_net_get_weights(self.h, _pf(w)), "get_weights")
        return w

    def forward(self, x, eb0):
        """y = N(x): x (nb, V, 3), eb0 (nb, V) node log-moduli -> (nb, V, 3)"""
        x = np.ascontiguousarray(x, dtype=np.float32)
        eb0 = np.ascontiguousarray(eb0, dtype=np.float32)
        nb = x.shape[0]
        y = np.zeros((nb, self.V, 3), dtype=np.float32)
        self._chk(lib().ufem_net_forward(self.h, nb, _pf(x), _pf(eb0), _pf(y)), "forward")
        return y

    def backward(self, x, eb0, dy):
        """gradient of sum <dy, N(x)> with respect to the packed parameters"""
        x = np.ascontiguousarray(x, dtype=np.float32)
        eb0 = np.ascontiguousarray(eb0, dtype=np.float32)
        dy = np.ascontiguousarray(dy, dtype=np.float32)
        g = np.zeros(self.P, dtype=np.float32)
        self._chk(lib().ufem_net_backward(self.h, x.shape[0], _pf(x), _pf(eb0), _pf(dy), _pf(g)), "backward")
        return g

    def pair(self, i):
        v = np.zeros((self.V, 4), dtype=np.float32)
        rr = np.zeros(3 * self.V, dtype=np.float32)
        ee = np.zeros(3 * self.V, dtype=np.float32)
        er = C.c_double(0)
        f = C.c_int(0)
        self._chk(lib().ufem_net_get_pair(self.h, i, _pf(v), _pf(rr), _pf(ee), C.byref(er), C.byref(f)), "pair")
        return v, rr, ee, er.value, f.value

    def stage_names(self):
        L = self.levels
        return ([f"C{l}" for l in range(L - 1)] + [f"C{L - 1}g"] +
                [f"C{l}T" for l in range(L - 1, -1, -1)])

    def apply64(self, v):
        """N(v) with the solver's FP64 inference, v [D][H][W][3] on the bound system's node grid"""
        v = np.ascontiguousarray(v, dtype=np.float64)
        y = np.zeros_like(v)
        rv = lib().ufem_net_apply64(self.h, _pd(v), _pd(y))
        if rv != AF_OK:
            raise RuntimeError(f"apply64 failed ({rv}): {last_error()}")
        return y

    def time_stages(self, reps=50):
        names = self.stage_names()
        us = np.zeros(len(names))
        self._chk(lib().ufem_net_time_stages(self.h, reps, _pd(us)), "time_stages")
        return dict(zip(names, us))

    def time_stages64(self, reps=50):
        """device us of each FP64 inference stage (the solver's path)"""
        names = self.stage_names()
        us = np.zeros(len(names))
        self._chk(lib().ufem_net_time_stages64(self.h, reps, _pd(us)), "time_stages64")
        return dict(zip(names, us))

    def eval(self, i0, cnt):
        out = np.zeros(cnt)
        self._chk(lib().ufem_net_eval(self.h, i0, cnt, _pd(out)), "eval")
        return out

    def bind(self, system, E):
        self._E = np.ascontiguousarray(E, dtype=np.float64)
        self._sys = system
        self._chk(lib().ufem_net_bind(self.h, system.h, _pd(self._E)), "bind")

    def reserve(self, cap):
        self._chk(lib().ufem_net_reserve(self.h, cap), "reserve")

    def npairs(self):
        return lib().ufem_net_npairs(self.h)

    def collect(self, pairs, eps=1e-11, every=8):
        info = SolveInfo()
        n = lib().ufem_net_collect(self.h, pairs, eps, every, C.byref(info))
        if n < 0:
            raise RuntimeError(f"collect failed ({n}): {last_error()}")
        return n, info.as_dict()

    def train(self, steps, batch=4, lr=1e-3, seed=0):
        loss = np.zeros(steps)
        ms = C.c_double(0)
        self._chk(lib().ufem_net_train(self.h, steps, batch, lr, seed, _pd(loss), C.byref(ms)), "train")
        return loss, ms.value


def probe_conv(mode, flip, x, w, gi, go, dy=None, x_f32=False):
    """One convolution kernel on host data: x (nb, Vi, cin), w (27, cout, cin)
    -> y (nb, Vo, cout) [and dw (27, cout, cin) for dy (nb, Vo, cout)].
    x_f32: False/0 bf16 activations, True/1 f32, 2 f32 stored split."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    nb, _, cin = x.shape
    cout = w.shape[1]
    Vo = int(np.prod(go))
    gi_ = np.ascontiguousarray(gi, dtype=np.int32)
    go_ = np.ascontiguousarray(go, dtype=np.int32)
    y = np.zeros((nb, Vo, cout), dtype=np.float32)
    dw = None
    if dy is not None:
        dy = np.ascontiguousarray(dy, dtype=np.float32)
        dw = np.zeros((27, cout, cin), dtype=np.float32)
    xf = 2 if (x_f32 is not True and x_f32 == 2) else (1 if x_f32 else 0)
    rv = lib().ufem_net_probe_conv(mode, flip, xf, nb, _pi(gi_), _pi(go_), cin, cout, _pf(x), _pf(w),
                                   _pf(y), _pf(dy) if dy is not None else None, _pf(dw) if dw is not None else None)
    if rv != 0:
        raise RuntimeError(f"probe_conv failed ({rv}): {last_error()}")
    return (y, dw) if dy is not None else y


def time_conv(mode, flip, x_f32, nb, gi, go, cin, cout, reps=50):
    gi_ = np.ascontiguousarray(gi, dtype=np.int32)
    go_ = np.ascontiguousarray(go, dtype=np.int32)
    us = C.c_double(0)
    rv = lib().ufem_net_time_conv(mode, flip, x_f32, nb, _pi(gi_), _pi(go_), cin, cout, reps, C.byref(us))
    if rv != 0:
        raise RuntimeError(f"time_conv failed ({rv}): {last_error()}")
    return us.value


def _mesh_struct(mesh, keep):
    arrs = [np.ascontiguousarray(mesh.coords[:, i], dtype=np.float64) for i in range(3)]
    arrs += [np.ascontiguousarray(mesh.enodes.reshape(-1), dtype=np.int64),
             np.ascontiguousarray(mesh.E, dtype=np.float64),
             np.ascontiguousarray(mesh.nu, dtype=np.float64),
             np.ascontiguousarray(mesh.nl_node, dtype=np.int64),
             np.ascontiguousarray(mesh.nl_type, dtype=np.int64),
             np.ascontiguousarray(mesh.nl_dir, dtype=np.int64),
             np.ascontiguousarray(mesh.nl_val, dtype=np.float64)]
    keep.append(arrs)
    k = arrs
    return _Mesh(mesh.n_nodes, _pd(k[0]), _pd(k[1]), _pd(k[2]), mesh.n_elems, _pl(k[3]), _pd(k[4]),
                 _pd(k[5]), k[6].size, _pl(k[6]), _pl(k[7]), _pl(k[8]), _pd(k[9]))


class GpuBatch:
    """Many samples solved in one persistent launch (ufem_batch_*)."""

    def __init__(self, meshes):
        keep = []
        arr = (_Mesh * len(meshes))(*[_mesh_struct(m, keep) for m in meshes])
        self.h = lib().ufem_batch_create(len(meshes), arr)
        if not self.h:
            raise RuntimeError("ufem_batch_create: " + last_error())
        self.n = len(meshes)
        self.nclasses = lib().ufem_batch_nclasses(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().ufem_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def assemble(self, npts=2):
        ms = C.c_double(0)
        rv = lib().ufem_batch_assemble(self.h, npts, C.byref(ms))
        if rv != 0:
            raise RuntimeError(f"batch assemble failed ({rv}): {last_error()}")
        return ms.value

    def autotune(self, cache=None, budget_iters=100):
        wg, grid = C.c_int(0), C.c_int(0)
        rv = lib().ufem_batch_autotune(self.h, cache.encode() if cache else None, budget_iters,
                                       C.byref(wg), C.byref(grid))
        if rv != 0:
            raise RuntimeError(f"autotune failed ({rv}): {last_error()}")
        return wg.value, grid.value

    def solve(self, eps=1e-8, maxit=0):
        n = self.n
        ms = C.c_double(0)
        it = np.zeros(n, dtype=np.int64)
        rr = np.zeros(n)
        st = np.zeros(n, dtype=np.int32)
        sm = np.zeros(n)
        rv = lib().ufem_batch_solve(self.h, eps, maxit, C.byref(ms), _pl(it), _pd(rr), _pi(st), _pd(sm))
        wg, grid, tb = C.c_int(0), C.c_int(0), C.c_int(0)
        lib().ufem_batch_shape(self.h, C.byref(wg), C.byref(grid), C.byref(tb))
        team = np.zeros(n, dtype=np.int32)
        lib().ufem_batch_teams(self.h, _pi(team))
        return {"rv": rv, "total_ms": ms.value, "iters": it, "relres": rr, "status": st, "sample_ms": sm,
                "wg": wg.value, "grid": grid.value, "team_big": tb.value, "team": team}

    def set_limits(self, wait_ms=0.0, timeout_ms=0.0):
        """team waits give up after wait_ms; the host stops a solve after timeout_ms"""
        lib().ufem_batch_set_limits(self.h, wait_ms, timeout_ms)

    def set_shape(self, wg, grid, team_big):
        lib().ufem_batch_set_shape(self.h, wg, grid, team_big)

    def set_mixed(self, mode=1):
        """FP64 vectors over FP32 copies of the samples' K (mode 1, True), with
        bf16 remainders beside the FP32 values (mode 2, 48-bit values) and FP64
        residual refreshes; 0 or False: K in FP64"""
        rv = lib().ufem_batch_set_mixed(self.h, int(mode))
        if rv != 0:
            raise ValueError(last_error())

    def u(self, i):
        out = np.zeros(lib().ufem_batch_ndof(self.h, i))
        rv = lib().ufem_batch_get_u(self.h, i, _pd(out))
        if rv != 0:
            raise RuntimeError(last_error())
        return out


def element_stiffness(xyz, E, nu, npts=2):
    """Batched element stiffness on the GPU: xyz (n, 8, 3) -> (n, 24, 24)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 24)
    n = xyz.shape[0]
    E = np.ascontiguousarray(np.broadcast_to(E, (n,)), dtype=np.float64)
    nu = np.ascontiguousarray(np.broadcast_to(nu, (n,)), dtype=np.float64)
    Ke = np.zeros((n, 576))
    rv = lib().ufem_gpu_element_stiffness(n, _pd(xyz), _pd(E), _pd(nu), npts, _pd(Ke))
    if rv != 0:
        raise RuntimeError(f"element stiffness failed ({rv}): {last_error()}")
    return Ke.reshape(n, 24, 24)


class RowSlotMatrix:
    """A uFEM row-slot matrix (fem/fem_math.h:47-57) owned by numpy arrays,
    for calling the drop-in entry points exactly as uFEM does."""

    @classmethod
    def empty(cls, n, capacity):
        """all slots empty (pos 0), as femMatAlloc leaves them (fem/fem_math.c:57-133)"""
        cap = np.asarray(capacity, dtype=np.int64)
        return cls(n, np.zeros(n, dtype=np.int64), np.zeros(0, dtype=np.int64), np.zeros(0), capacity=cap)

    def rows_csr(self):
        """(indptr, 0-based cols, vals) of the filled slots, each row sorted by column"""
        n = self.s.rows
        rl = np.zeros(n, dtype=np.int64)
        for i in range(n):
            f, c = int(self.frompos[i]), int(self.defpos[i])
            seg = self.pos[f:f + c]
            z = np.nonzero(seg <= 0)[0]
            rl[i] = z[0] if z.size else c
        ip = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(rl, out=ip[1:])
        cols = np.empty(ip[-1], dtype=np.int64)
        vals = np.empty(ip[-1])
        for i in range(n):
            f, k = int(self.frompos[i]), int(rl[i])
            o = np.argsort(self.pos[f:f + k], kind="stable")
            cols[ip[i]:ip[i + 1]] = self.pos[f:f + k][o] - 1
            vals[ip[i]:ip[i + 1]] = self.data[f:f + k][o]
        return ip, cols, vals

    def mirror_info(self):
        """(block size, rows, filled slots, build ms, refresh ms) of its device copy"""
        out = np.zeros(5)
        rv = lib().ufem_gpu_mirror_info(C.byref(self.s), _pd(out))
        if rv != AF_OK:
            raise RuntimeError(last_error())
        return {"bs": int(out[0]), "rows": int(out[1]), "nfill": int(out[2]), "build_ms": out[3],
                "refresh_ms": out[4]}

    def mark_dirty(self):
        lib().ufem_gpu_mark_dirty(C.byref(self.s))

    def release(self):
        lib().ufem_gpu_release(C.byref(self.s))

    def set_zero(self):
        lib().femMatSetZeroBig(C.byref(self.s))

    def __init__(self, n, rowlen, cols1, vals, capacity=None):
        rowlen = np.asarray(rowlen, dtype=np.int64)
        cap = rowlen if capacity is None else np.asarray(capacity, dtype=np.int64)
        self.defpos = np.ascontiguousarray(cap)
        self.frompos = np.zeros(n, dtype=np.int64)
        np.cumsum(cap[:-1], out=self.frompos[1:])
        ln = int(cap.sum())
        self.pos = np.zeros(ln + 1, dtype=np.int64)
        self.data = np.zeros(ln + 1)
        tot = int(rowlen.sum())
        if tot:
            # slot of entry j of row i: frompos[i] + j
            start = np.zeros(n, dtype=np.int64)
            np.cumsum(rowlen[:-1], out=start[1:])
            row = np.repeat(np.arange(n, dtype=np.int64), rowlen)
            at = self.frompos[row] + (np.arange(tot, dtype=np.int64) - start[row])
            self.pos[at] = np.asarray(cols1, dtype=np.int64)[:tot]
            self.data[at] = np.asarray(vals, dtype=np.float64)[:tot]
        self.s = tMatrix(1, n, n, ln, _pl(self.pos), _pd(self.data), _pl(self.frompos), _pl(self.defpos))


class Vec:
    def __init__(self, data):
        self.data = np.ascontiguousarray(data, dtype=np.float64).copy()
        self.s = tVector(0, self.data.size, self.data.size, None, _pd(self.data))


def femEqsCGwJ(K: RowSlotMatrix, b: Vec, x: Vec, eps: float, maxit: int) -> int:
    return lib().femEqsCGwJ(C.byref(K.s), C.byref(b.s), C.byref(x.s), eps, maxit)


def femEqsBiCCSwJ(K: RowSlotMatrix, b: Vec, x: Vec, eps: float, maxit: int) -> int:
    return lib().femEqsBiCCSwJ(C.byref(K.s), C.byref(b.s), C.byref(x.s), eps, maxit)


def femEqsCGwSSOR(K: RowSlotMatrix, b: Vec, x: Vec, eps: float, maxit: int) -> int:
    return lib().femEqsCGwSSOR(C.byref(K.s), C.byref(b.s), C.byref(x.s), eps, maxit)


def set_stop_rule(rule: int) -> None:
    lib().ufem_gpu_set_stop_rule(rule)


def set_mirror_policy(policy: int) -> None:
    lib().ufem_gpu_set_mirror_policy(policy)


def femMatVecMultBig(K: RowSlotMatrix, b: Vec, c: Vec) -> int:
    return lib().femMatVecMultBig(C.byref(K.s), C.byref(b.s), C.byref(c.s))


def fill_K(K: RowSlotMatrix, mesh, dof_map, known_dofs=7, npts=3) -> int:
    """fem_fill_K(AF_NO) on the GPU into the row-slot K through the DOF map
    (nDOFfld: known_dofs entries per node, 1-based positions, 0 = none)"""
    keep = []
    m = _mesh_struct(mesh, keep)
    dm = np.ascontiguousarray(dof_map, dtype=np.int64)
    return lib().ufem_gpu_fill_K(C.byref(K.s), C.byref(m), _pl(dm), known_dofs, npts)


def fill_K_timings():
    t = np.zeros(3)
    lib().ufem_gpu_fill_K_timings(_pd(t))
    return {"setup_ms": t[0], "device_ms": t[1], "writeback_ms": t[2]}


def dof_map(mesh, fbc=False, known_dofs=7):
    """the reference's nDOFfld (fem_dofs, fem/fem_sol.c:1263-1391): node-major
    1-based positions of ux, uy, uz; with fbc the DOFs of zero-valued supports
    are left out (femFastBC, :1280-1296)"""
    nn = mesh.n_nodes
    drop = np.zeros(nn * known_dofs, dtype=bool)
    if fbc:
        for nd, t, d, v in zip(mesh.nl_node, mesh.nl_type, mesh.nl_dir, mesh.nl_val):
            if int(t) == 1 and abs(float(v)) <= 1e-17:
                drop[int(nd) * known_dofs + int(d) - 1] = True
    act = np.zeros(nn * known_dofs, dtype=bool)
    used = np.zeros(nn, dtype=bool)
    used[np.asarray(mesh.enodes).ravel()] = True
    for d in range(3):
        act[np.nonzero(used)[0] * known_dofs + d] = True
    act &= ~drop
    m = np.zeros(nn * known_dofs, dtype=np.int64)
    m[act] = np.arange(1, int(act.sum()) + 1)
    return m


def last_solve() -> dict:
    info = SolveInfo()
    lib().ufem_gpu_last_solve(C.byref(info))
    return info.as_dict()


class Monte:
    """The Monte plugin interface (include/ufem_monte.h; fem/fem_mont.c): one
    model per process, read from a preparsed input file with a
    random-variable section; each solution() is one sample on the GPU."""

    def __init__(self, path: str):
        self.param = os.fsencode(path)
        rv = lib().monte_init_lib_stuff(self.param)
        if rv != 0:
            raise RuntimeError(f"monte_init_lib_stuff({path}) = {rv}: {last_error()}")
        il, ol, ff = C.c_long(), C.c_long(), C.c_long()
        lib().monte_nums_of_vars(self.param, C.byref(il), C.byref(ol), C.byref(ff))
        self.ilen, self.olen, self.ffunc = il.value, ol.value, ff.value

    @staticmethod
    def interface_type() -> int:
        return lib().monte_dlib_interface_type()

    def inames(self):
        return [lib().monte_ivar_name(self.param, i).decode() for i in range(self.ilen)]

    def onames(self):
        return [lib().monte_ovar_name(self.param, i).decode() for i in range(self.olen)]

    def solution(self, ifld, if_type: int = 0, ofld_len: int | None = None):
        """(return code, output field); the field has olen entries unless the
        input file places outputs further out"""
        x = np.ascontiguousarray(ifld, dtype=np.float64)
        y = np.zeros(max(self.olen, ofld_len or 0, 1))
        rv = lib().monte_solution(self.param, _pd(x), _pd(y), if_type)
        return rv, y

    def solution_batch(self, ifld, if_type: int = 0):
        """n samples ([n, ilen]) in one batched launch: (return code,
        [n, olen] output fields, per-sample status)"""
        x = np.ascontiguousarray(ifld, dtype=np.float64).reshape(-1, self.ilen)
        n = x.shape[0]
        y = np.zeros((n, max(self.olen, 1)))
        st = np.zeros(n, dtype=np.int32)
        rv = lib().monte_solution_batch(self.param, n, _pd(x), _pd(y), if_type, _pi(st))
        return rv, y, st

    def close(self):
        lib().monte_clean_lib_stuff(self.param)


class Model:
    """uFEM's input file read by the library (femReadInput for the brick
    path): preparsed (-p, positions) or with IDs renumbered to positions."""

    def __init__(self, path: str, preparsed: bool = True):
        rv = C.c_int(0)
        self.h = lib().ufem_model_read(os.fsencode(path), 1 if preparsed else 0, C.byref(rv))
        self.rv = rv.value
        if not self.h:
            raise RuntimeError(f"ufem_model_read({path}) = {rv.value}: {last_error()}")
        m = _Mesh()
        lib().ufem_model_mesh(self.h, C.byref(m))
        nn, ne, nl = m.n_nodes, m.n_elems, m.n_nl

        def arr(p, k, dt):
            return np.ctypeslib.as_array(p, shape=(k,)).astype(dt, copy=True) if k > 0 else np.zeros(0, dt)
        from .mesh import Mesh
        coords = np.stack([arr(m.x, nn, np.float64), arr(m.y, nn, np.float64), arr(m.z, nn, np.float64)], axis=1)
        self.mesh = Mesh(0, 0, 0, coords, arr(m.enodes, 8 * ne, np.int64).reshape(ne, 8), arr(m.E, ne, np.float64),
                         arr(m.nu, ne, np.float64), arr(m.nl_node, nl, np.int64), arr(m.nl_type, nl, np.int64),
                         arr(m.nl_dir, nl, np.int64), arr(m.nl_val, nl, np.float64))
        self.node_ids = arr(lib().ufem_model_node_ids(self.h), nn, np.int64)
        self.elem_ids = arr(lib().ufem_model_elem_ids(self.h), ne, np.int64)
        self.elem_mat = arr(lib().ufem_model_elem_mat(self.h), ne, np.int64)
        self.elem_rset = arr(lib().ufem_model_elem_rset(self.h), ne, np.int64)
        self.sol_id = lib().ufem_model_sol_id(self.h)
        self.sol_dep = lib().ufem_model_sol_dep(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().ufem_model_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_model(path: str, preparsed: bool = True) -> Model:
    return Model(path, preparsed)


def model_error(path: str, preparsed: bool = True):
    """the AF_* code and message of reading a file that fails"""
    rv = C.c_int(0)
    h = lib().ufem_model_read(os.fsencode(path), 1 if preparsed else 0, C.byref(rv))
    if h:
        lib().ufem_model_free(h)
        return 0, ""
    return rv.value, last_error()


def dense_spd_inverse(A):
    """inverse of a dense SPD matrix on the device (block Gauss-Jordan)"""
    A = np.ascontiguousarray(A, dtype=np.float64)
    out = np.zeros_like(A)
    rv = lib().ufem_gpu_dense_spd_inverse(A.shape[0], _pd(A), _pd(out))
    if rv != AF_OK:
        raise RuntimeError(f"dense inverse = {rv}: {last_error()}")
    return out
