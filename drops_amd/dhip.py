"""ctypes binding of libdhip (include/dhip.h).

The GPU library is the only compute path: if libdhip.so is missing or cannot
be loaded, every entry point raises -- there is no CPU fallback.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdhip.so")

OK, E_ZERO_NODE, E_ZERO_PATTERN, E_ARG, E_NOT_CONVERGED, E_NOMEM, E_HIP, E_STATE, E_CAPACITY, E_IO = range(10)
MTX_MARKET, MTX_PLAIN = 0, 1
HOST, DEVICE = 0, 1
ZERO_THROW, ZERO_ALLOW = 0, 1
OP_MASS, OP_LB, OP_STAB, OP_COMB = range(4)
RHS_NONE, RHS_3XY_MINUS_X3, RHS_CONST, RHS_VALUES = range(4)
PC_NONE, PC_JACOBI = 0, 1
FP64, FP32_IR, FP64_CLASSIC = 0, 1, 2
LOC_N = 57
LOC2_N = 320

# every symbol declared in include/dhip.h
EXPORTS = [
    "dhip_create", "dhip_destroy", "dhip_last_error", "dhip_profile_enable", "dhip_profile_get",
    "dhip_profile_reset", "dhip_solver_probe", "dhip_tune", "dhip_malloc", "dhip_free", "dhip_memcpy", "dhip_synchronize",
    "dhip_mesh_kuhn", "dhip_levelset_p2", "dhip_classify", "dhip_local_p1", "dhip_number_p1if",
    "dhip_build_p1", "dhip_assemble_p1", "dhip_get_cut_tets", "dhip_get_local", "dhip_get_tet_dofs",
    "dhip_get_dof_vertices", "dhip_get_csr", "dhip_device_arrays", "dhip_csr_create",
    "dhip_csr_from_assembly", "dhip_csr_destroy", "dhip_csr_info", "dhip_spmv", "dhip_dot",
    "dhip_axpy", "dhip_pcg", "dhip_gmres", "dhip_gmres_diag", "dhip_ch_create", "dhip_ch_destroy",
    "dhip_ch_set_state", "dhip_ch_get_state", "dhip_ch_prepare", "dhip_ch_system", "dhip_ch_solve", "dhip_ch_step",
    "dhip_mlp_create", "dhip_mlp_destroy", "dhip_mlp_params", "dhip_mlp_normalization", "dhip_mlp_features",
    "dhip_mlp_targets", "dhip_mlp_predict", "dhip_mlp_grad", "dhip_mlp_train_step", "dhip_ch_history",
    "dhip_ch_residual", "dhip_ch_predict_start", "dhip_ch_solve_learned", "dhip_local_p2", "dhip_number_p2if",
    "dhip_build_p2", "dhip_assemble_p2", "dhip_get_local_p2", "dhip_get_dof_nodes", "dhip_ch_create_p2",
    "dhip_batch_create", "dhip_batch_solve", "dhip_batch_stats", "dhip_batch_problem", "dhip_batch_destroy",
    "dhip_mtx_write", "dhip_mtx_format", "dhip_mtx_read", "dhip_mtx_parse", "dhip_csr_export", "dhip_csr_import",
    "dhip_vec_write", "dhip_vec_format", "dhip_vec_read", "dhip_vec_parse", "dhip_vec_count",
    "dhip_assemble_p1_band", "dhip_pcg_noalloc", "dhip_ch_bdf_step", "dhip_ch_set_precision", "dhip_spmv_t",
    "dhip_quad_points", "dhip_rhs_values", "dhip_local_p2_quaqua", "dhip_assemble_p2_quaqua",
    "dhip_p2_recovered_gradient", "dhip_quaqua_points", "dhip_quaqua_values",
]
MLP_NPARAM, MLP_NF = 9282, 12


class ChParams(C.Structure):
    """dhip_ch_params_t"""
    _fields_ = [("dt", C.c_double), ("epsilon", C.c_double), ("rho", C.c_double), ("sigma", C.c_double),
                ("S", C.c_double), ("restart", C.c_int32), ("max_iter", C.c_int32), ("tol", C.c_double),
                ("rel", C.c_int32)]


class QuaQuaParams(C.Structure):
    """dhip_quaqua_params_t: the QuaQuaMapperCL parameters (LevelsetMapper.Iter / Tol / ArmijoConstant,
    damping steps); the defaults are the reference's parameter files'."""
    _fields_ = [("maxiter", C.c_int32), ("tol", C.c_double), ("armijo_c", C.c_double), ("max_damping", C.c_int32)]

    def __init__(self, maxiter=20, tol=1e-7, armijo_c=1e-2, max_damping=8):
        super().__init__(maxiter, tol, armijo_c, max_damping)


class BdfParams(C.Structure):
    """dhip_bdf_params_t"""
    _fields_ = [("step_size", C.c_double), ("tol", C.c_double), ("rho", C.c_double), ("min_step", C.c_double),
                ("max_step", C.c_double), ("adaptive", C.c_int32), ("max_tries", C.c_int32)]


class BdfStat(C.Structure):
    """dhip_bdf_stat_t"""
    _fields_ = [("step", C.c_int64), ("t", C.c_double), ("dt", C.c_double), ("e", C.c_double),
                ("tries", C.c_int32), ("iters_bdf1", C.c_int32), ("iters_bdf2", C.c_int32),
                ("reserved", C.c_int32), ("resid", C.c_double)]


class BatchProblem(C.Structure):
    """dhip_batch_problem_t"""
    _fields_ = [("orig", C.c_double * 3), ("e", C.c_double * 9), ("n", C.c_int32 * 3), ("phi_offset", C.c_int64),
                ("lc", C.c_double * 3)]


class BatchStat(C.Structure):
    """dhip_batch_stat_t"""
    _fields_ = [("status", C.c_int32), ("ncut", C.c_int32), ("ndof", C.c_int32), ("nnz", C.c_int32),
                ("iters", C.c_int32), ("outer", C.c_int32), ("resid", C.c_double), ("start_s", C.c_double),
                ("end_s", C.c_double), ("assembled_s", C.c_double)]


class DhipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dhip error {code}: {msg}")
        self.code = code


_lib = None


def load(path=LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(path)
    P, D, I32, I64, VP = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.c_void_p
    L.dhip_create.argtypes = [C.POINTER(P), C.c_int32]
    L.dhip_destroy.argtypes = [P]
    L.dhip_destroy.restype = None
    L.dhip_last_error.argtypes = [P]
    L.dhip_last_error.restype = C.c_char_p
    L.dhip_profile_enable.argtypes = [P, C.c_int32]
    L.dhip_profile_get.argtypes = [P, C.c_char_p, D, I64]
    L.dhip_profile_reset.argtypes = [P]
    L.dhip_solver_probe.argtypes = [P, C.c_int32, C.POINTER(C.c_uint64), I32]
    L.dhip_tune.argtypes = [P, C.c_char_p, C.c_int64]
    L.dhip_malloc.argtypes = [P, C.POINTER(VP), C.c_int64]
    L.dhip_free.argtypes = [P, VP]
    L.dhip_memcpy.argtypes = [P, VP, VP, C.c_int64, C.c_int32]
    L.dhip_synchronize.argtypes = [P]
    L.dhip_mesh_kuhn.argtypes = [P, D, D, I32]
    L.dhip_levelset_p2.argtypes = [P, VP, C.c_int64, C.c_int32]
    L.dhip_classify.argtypes = [P, C.c_int32, I64]
    L.dhip_local_p1.argtypes = [P, D, C.c_int32, D]
    L.dhip_number_p1if.argtypes = [P, C.c_double, I64]
    L.dhip_build_p1.argtypes = [P, D, I64]
    L.dhip_assemble_p1.argtypes = [P, C.c_int32, D, C.c_int32, D, C.c_double, D, I64, I64, I64]
    L.dhip_get_cut_tets.argtypes = [P, I64]
    L.dhip_get_local.argtypes = [P, D, C.POINTER(C.c_uint8)]
    L.dhip_local_p2.argtypes = [P, D, C.c_int32, D]
    L.dhip_number_p2if.argtypes = [P, C.c_double, I64]
    L.dhip_build_p2.argtypes = [P, D, I64]
    L.dhip_assemble_p2.argtypes = [P, C.c_int32, D, C.c_int32, D, C.c_double, D, I64, I64, I64]
    L.dhip_local_p2_quaqua.argtypes = [P, C.POINTER(QuaQuaParams), C.c_int32, D]
    L.dhip_p2_recovered_gradient.argtypes = [P, D, C.c_int32]
    L.dhip_quaqua_points.argtypes = [P, C.POINTER(QuaQuaParams), I64, D, C.c_int32]
    L.dhip_quaqua_values.argtypes = [P, D, C.c_int64, C.c_int32]
    L.dhip_assemble_p2_quaqua.argtypes = [P, C.c_int32, C.POINTER(QuaQuaParams), C.c_int32, D, C.c_double, D, I64,
                                          I64, I64]
    L.dhip_assemble_p1_band.argtypes = [P, C.c_int32, D, C.c_int32, D, C.c_double, D, I64, I64, I64]
    L.dhip_get_local_p2.argtypes = [P, D, C.POINTER(C.c_uint8)]
    L.dhip_get_dof_nodes.argtypes = [P, I64]
    L.dhip_get_tet_dofs.argtypes = [P, I32]
    L.dhip_get_dof_vertices.argtypes = [P, I64]
    L.dhip_get_csr.argtypes = [P, I32, I32, D, D, D, D, D]
    L.dhip_device_arrays.argtypes = [P, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP), C.POINTER(VP)]
    L.dhip_csr_create.argtypes = [P, C.c_int64, C.c_int64, VP, VP, VP, C.c_int32, C.POINTER(P)]
    L.dhip_csr_from_assembly.argtypes = [P, C.c_int32, C.POINTER(P)]
    L.dhip_csr_destroy.argtypes = [P]
    L.dhip_csr_destroy.restype = None
    L.dhip_csr_info.argtypes = [P, I64, I64]
    L.dhip_spmv.argtypes = [P, P, VP, VP, C.c_int32]
    L.dhip_spmv_t.argtypes = [P, P, C.c_int64, VP, VP, C.c_int32]
    L.dhip_quad_points.argtypes = [P, I64, VP, C.c_int32]
    L.dhip_rhs_values.argtypes = [P, VP, C.c_int64, C.c_int32]
    L.dhip_dot.argtypes = [P, C.c_int64, VP, VP, D, C.c_int32]
    L.dhip_axpy.argtypes = [P, C.c_int64, C.c_double, VP, VP, C.c_int32]
    L.dhip_pcg.argtypes = [P, P, C.c_int32, VP, VP, C.c_int32, I32, D, C.c_int32, C.c_int32]
    L.dhip_pcg_noalloc.argtypes = [P, P, C.c_int32, VP, VP, I32, D, C.c_int32]
    L.dhip_gmres.argtypes = [P, P, C.c_int32, VP, VP, C.c_int32, C.c_int32, I32, D, C.c_int32, C.c_int32]
    L.dhip_gmres_diag.argtypes = [P, P, VP, VP, VP, C.c_int32, C.c_int32, I32, D, C.c_int32, C.c_int32]
    L.dhip_ch_create.argtypes = [P, C.c_int32, C.POINTER(ChParams), C.POINTER(P), I64]
    L.dhip_ch_create_p2.argtypes = [P, C.c_int32, C.POINTER(ChParams), C.POINTER(P), I64]
    L.dhip_ch_destroy.argtypes = [P]
    L.dhip_ch_destroy.restype = None
    L.dhip_ch_set_state.argtypes = [P, VP, VP, C.c_int32]
    L.dhip_ch_get_state.argtypes = [P, VP, VP, C.c_int32]
    L.dhip_ch_prepare.argtypes = [P]
    L.dhip_ch_system.argtypes = [P, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP), C.POINTER(VP), C.POINTER(VP),
                                 C.POINTER(VP), I64, I64]
    L.dhip_ch_solve.argtypes = [P, VP, C.c_int32, I32, D]
    L.dhip_ch_step.argtypes = [P, C.c_int32, I32, D]
    L.dhip_ch_bdf_step.argtypes = [P, C.POINTER(BdfParams), C.c_int32, C.POINTER(BdfStat)]
    L.dhip_ch_set_precision.argtypes = [P, C.c_int32]
    F = C.POINTER(C.c_float)
    L.dhip_mlp_create.argtypes = [P, C.POINTER(P)]
    L.dhip_mlp_destroy.argtypes = [P]
    L.dhip_mlp_destroy.restype = None
    L.dhip_mlp_params.argtypes = [P, F, F]
    L.dhip_mlp_normalization.argtypes = [P, D, D, D, D]
    L.dhip_mlp_features.argtypes = [P, C.POINTER(VP), C.c_int32, VP, VP, C.c_int64, VP]
    L.dhip_mlp_targets.argtypes = [P, VP, C.POINTER(VP), C.c_int32, C.c_int64, VP]
    L.dhip_mlp_predict.argtypes = [P, VP, C.c_int64, VP]
    L.dhip_mlp_grad.argtypes = [P, VP, VP, VP, C.c_int64, D, D]
    L.dhip_mlp_train_step.argtypes = [P, VP, VP, VP, C.c_int64, C.c_float, D]
    L.dhip_ch_history.argtypes = [P, C.c_int32, C.POINTER(VP)]
    L.dhip_ch_residual.argtypes = [P, VP]
    L.dhip_ch_predict_start.argtypes = [P, P, VP, C.c_int32, I32]
    L.dhip_ch_solve_learned.argtypes = [P, P, I32, D, I32]
    L.dhip_batch_create.argtypes = [P, C.c_int32, C.POINTER(BatchProblem), VP, C.c_int64, C.c_int32, C.c_int32, D,
                                    C.c_int32, D, C.c_double, C.POINTER(P)]
    L.dhip_batch_solve.argtypes = [P, C.c_int32, C.c_double, C.c_int32, C.c_int32]
    L.dhip_batch_stats.argtypes = [P, C.POINTER(BatchStat)]
    L.dhip_batch_problem.argtypes = [P, C.c_int32, I32, I32, D, D, D]
    L.dhip_batch_destroy.argtypes = [P]
    L.dhip_batch_destroy.restype = None
    CP = C.c_char_p
    L.dhip_mtx_write.argtypes = [CP, C.c_int32, C.c_int64, C.c_int64, I32, I32, D]
    L.dhip_mtx_read.argtypes = [CP, C.c_int32, I64, I64, I64, I32, I32, D]
    L.dhip_csr_export.argtypes = [P, CP, C.c_int32]
    L.dhip_csr_import.argtypes = [P, CP, C.c_int32, C.POINTER(P)]
    L.dhip_vec_write.argtypes = [P, CP, VP, C.c_int64, C.c_int32, C.c_int32]
    L.dhip_vec_read.argtypes = [P, CP, VP, C.c_int64, C.c_int32, C.c_int32]
    L.dhip_vec_count.argtypes = [CP, C.c_int32, I64]
    L.dhip_mtx_format.argtypes = [C.c_int32, C.c_int64, C.c_int64, I32, I32, D, VP, C.c_int64, I64]
    L.dhip_mtx_parse.argtypes = [VP, C.c_int64, C.c_int32, I64, I64, I64, I32, I32, D]
    L.dhip_vec_format.argtypes = [P, VP, C.c_int64, C.c_int32, C.c_int32, VP, C.c_int64, I64]
    L.dhip_vec_parse.argtypes = [P, VP, C.c_int64, VP, C.c_int64, C.c_int32, C.c_int32]
    _lib = L
    return L


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def _vp(a):
    if a is None:
        return None
    if isinstance(a, int):
        return C.c_void_p(a)
    return C.c_void_p(a.ctypes.data)


# ---- on-disk formats on host memory (no GPU needed) ----
def _fmt(fmt):
    return {"mtx": MTX_MARKET, "plain": MTX_PLAIN}[fmt] if isinstance(fmt, str) else int(fmt)


def _check_noctx(L, rc):
    if rc != OK:
        raise DhipError(rc, L.dhip_last_error(None).decode())


def mtx_write(path, rowptr, colind, val, ncols=None, fmt="mtx"):
    """MatrixCL::exportMTX (fmt "mtx") or out() (fmt "plain") of a host CSR matrix."""
    L = load()
    rp = np.ascontiguousarray(rowptr, np.int32)
    ci = np.ascontiguousarray(colind, np.int32)
    v = np.ascontiguousarray(val, np.float64)
    n = rp.size - 1
    _check_noctx(L, L.dhip_mtx_write(os.fsencode(path), _fmt(fmt), n, n if ncols is None else int(ncols),
                                     _p(rp, C.c_int32), _p(ci, C.c_int32), _p(v, C.c_double)))


def mtx_read(path, fmt="mtx"):
    """dict(shape, rowptr, colind, val) of a matrix file (MatrixMarket or plain)."""
    L = load()
    nr, nc, nz = C.c_int64(), C.c_int64(), C.c_int64()
    b = os.fsencode(path)
    _check_noctx(L, L.dhip_mtx_read(b, _fmt(fmt), C.byref(nr), C.byref(nc), C.byref(nz), None, None, None))
    rp = np.zeros(nr.value + 1, np.int32)
    ci = np.zeros(max(nz.value, 1), np.int32)
    v = np.zeros(max(nz.value, 1), np.float64)
    _check_noctx(L, L.dhip_mtx_read(b, _fmt(fmt), C.byref(nr), C.byref(nc), C.byref(nz), _p(rp, C.c_int32),
                                    _p(ci, C.c_int32), _p(v, C.c_double)))
    return dict(shape=(nr.value, nc.value), rowptr=rp, colind=ci[:nz.value], val=v[:nz.value])


def mtx_format(rowptr, colind, val, ncols=None, fmt="mtx"):
    """bytes of operator<< (fmt "mtx") or out() (fmt "plain") of a host CSR matrix"""
    L = load()
    rp = np.ascontiguousarray(rowptr, np.int32)
    ci = np.ascontiguousarray(colind, np.int32)
    v = np.ascontiguousarray(val, np.float64)
    n = rp.size - 1
    nc = n if ncols is None else int(ncols)
    cap = C.c_int64()
    args = (_fmt(fmt), n, nc, _p(rp, C.c_int32), _p(ci, C.c_int32), _p(v, C.c_double))
    _check_noctx(L, L.dhip_mtx_format(*args, None, 0, C.byref(cap)))
    buf = C.create_string_buffer(max(cap.value, 1))
    ln = C.c_int64()
    _check_noctx(L, L.dhip_mtx_format(*args, C.cast(buf, C.c_void_p), cap.value, C.byref(ln)))
    return buf.raw[:ln.value]


def mtx_parse(data, fmt="mtx"):
    """dict(shape, rowptr, colind, val) of matrix text held in memory"""
    L = load()
    b = bytes(data)
    nr, nc, nz = C.c_int64(), C.c_int64(), C.c_int64()
    _check_noctx(L, L.dhip_mtx_parse(b, len(b), _fmt(fmt), C.byref(nr), C.byref(nc), C.byref(nz), None, None, None))
    rp = np.zeros(nr.value + 1, np.int32)
    ci = np.zeros(max(nz.value, 1), np.int32)
    v = np.zeros(max(nz.value, 1), np.float64)
    _check_noctx(L, L.dhip_mtx_parse(b, len(b), _fmt(fmt), C.byref(nr), C.byref(nc), C.byref(nz),
                                     _p(rp, C.c_int32), _p(ci, C.c_int32), _p(v, C.c_double)))
    return dict(shape=(nr.value, nc.value), rowptr=rp, colind=ci[:nz.value], val=v[:nz.value])


def vec_format(x, binary=True):
    """bytes of VecDescBaseCL::Write of a host vector"""
    L = load()
    a = np.ascontiguousarray(x, np.float64)
    cap = C.c_int64()
    _check_noctx(L, L.dhip_vec_format(None, _vp(a), a.size, HOST, int(binary), None, 0, C.byref(cap)))
    buf = C.create_string_buffer(max(cap.value, 1))
    ln = C.c_int64()
    _check_noctx(L, L.dhip_vec_format(None, _vp(a), a.size, HOST, int(binary), C.cast(buf, C.c_void_p), cap.value,
                                      C.byref(ln)))
    return buf.raw[:ln.value]


def vec_parse(data, n, binary=True):
    """VecDescBaseCL::Read of n values from bytes held in memory"""
    L = load()
    b = bytes(data)
    out = np.zeros(max(int(n), 1), np.float64)
    _check_noctx(L, L.dhip_vec_parse(None, b, len(b), _vp(out), int(n), HOST, int(binary)))
    return out[:int(n)]


def vec_write(path, x, binary=True):
    """VecDescBaseCL::Write of a host vector."""
    L = load()
    a = np.ascontiguousarray(x, np.float64)
    _check_noctx(L, L.dhip_vec_write(None, os.fsencode(path), _vp(a), a.size, HOST, int(binary)))


def vec_count(path, binary=True):
    L = load()
    n = C.c_int64()
    _check_noctx(L, L.dhip_vec_count(os.fsencode(path), int(binary), C.byref(n)))
    return n.value


def vec_read(path, n=None, binary=True):
    """VecDescBaseCL::Read into a new host vector; n (default: the stored count) must match."""
    L = load()
    n = vec_count(path, binary) if n is None else int(n)
    out = np.zeros(max(n, 1), np.float64)
    _check_noctx(L, L.dhip_vec_read(None, os.fsencode(path), _vp(out), n, HOST, int(binary)))
    return out[:n]


class Csr:
    """Device CSR matrix handle (MatrixCL mirror).  Owned by its Context: closing the
    context closes its matrices first (a handle must not outlive its context)."""

    def __init__(self, ctx, handle):
        self.ctx, self.h = ctx, handle
        ctx._csrs.add(self)

    @property
    def shape(self):
        n, z = C.c_int64(), C.c_int64()
        self.ctx.L.dhip_csr_info(self.h, C.byref(n), C.byref(z))
        return n.value, z.value

    def export(self, path, fmt="mtx"):
        """MatrixCL::exportMTX (or out()) of the device matrix."""
        self.ctx.check(self.ctx.L.dhip_csr_export(self.h, os.fsencode(path), _fmt(fmt)))

    def close(self):
        if self.h and self.ctx.h:
            self.ctx.L.dhip_csr_destroy(self.h)
        self.h = None
        self.ctx._csrs.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CahnHilliard:
    """Device Cahn-Hilliard stepper (CahnHilliardcGP1CL mirror, theta = 1, resting interface).
    Unknowns [mu; c] on the P1 (space=1) or P2 (space=2) interface space of the context's mesh
    and level set."""

    def __init__(self, ctx, dt, epsilon, rho, sigma=1.0, S=1.0, restart=50, max_iter=20000, tol=1e-8, rel=True,
                 zero_policy=ZERO_THROW, space=1):
        self.ctx, self.L = ctx, ctx.L
        self.space = space
        self.params = ChParams(dt, epsilon, rho, sigma, S, restart, max_iter, tol, int(rel))
        self.h = C.c_void_p()
        n = C.c_int64()
        create = self.L.dhip_ch_create_p2 if space == 2 else self.L.dhip_ch_create
        ctx.check(create(ctx.h, zero_policy, C.byref(self.params), C.byref(self.h), C.byref(n)))
        self.n = n.value
        ctx._csrs.add(self)

    def set_state(self, c, mu):
        c = np.ascontiguousarray(c, np.float64)
        mu = np.ascontiguousarray(mu, np.float64)
        assert c.size == self.n and mu.size == self.n
        self.ctx.check(self.L.dhip_ch_set_state(self.h, _vp(c), _vp(mu), HOST))

    def state(self):
        c, mu = np.zeros(self.n), np.zeros(self.n)
        self.ctx.check(self.L.dhip_ch_get_state(self.h, _vp(c), _vp(mu), HOST))
        return c, mu

    def prepare(self):
        self.ctx.check(self.L.dhip_ch_prepare(self.h))

    def system(self):
        """Device pointers of the step system and the state vector."""
        ptr = [C.c_void_p() for _ in range(6)]
        n2, z2 = C.c_int64(), C.c_int64()
        self.ctx.check(self.L.dhip_ch_system(self.h, *[C.byref(p) for p in ptr], C.byref(n2), C.byref(z2)))
        keys = ("rowptr", "col", "val", "rhs", "pcdiag", "state")
        out = {k: p.value for k, p in zip(keys, ptr)}
        out.update(n2=n2.value, nnz2=z2.value)
        return out

    def system_host(self):
        """Host copies of the step system (block CSR, rhs, preconditioner diagonal, state)."""
        s = self.system()
        ctx = self.ctx
        return dict(rowptr=ctx.to_host(s["rowptr"], s["n2"] + 1, np.int32), col=ctx.to_host(s["col"], s["nnz2"], np.int32),
                    val=ctx.to_host(s["val"], s["nnz2"], np.float64), rhs=ctx.to_host(s["rhs"], s["n2"], np.float64),
                    pcdiag=ctx.to_host(s["pcdiag"], s["n2"], np.float64),
                    state=ctx.to_host(s["state"], s["n2"], np.float64))

    def solve(self, x0=None, where=HOST, raise_on_fail=True):
        """x0: [mu; c] start (host array, or device pointer with where=DEVICE); None = previous step."""
        it, t = C.c_int32(), C.c_double()
        if x0 is None:
            rc = self.L.dhip_ch_solve(self.h, None, HOST, C.byref(it), C.byref(t))
        elif where == DEVICE:
            rc = self.L.dhip_ch_solve(self.h, C.c_void_p(x0), DEVICE, C.byref(it), C.byref(t))
        else:
            x0 = np.ascontiguousarray(x0, np.float64)
            assert x0.size == 2 * self.n
            rc = self.L.dhip_ch_solve(self.h, _vp(x0), HOST, C.byref(it), C.byref(t))
        if rc not in (OK, E_NOT_CONVERGED) or (raise_on_fail and rc != OK):
            self.ctx.check(rc)
        return it.value, t.value, rc == OK

    def history(self, level):
        """Device pointer of the state `level` steps back (0 = current)."""
        p = C.c_void_p()
        self.ctx.check(self.L.dhip_ch_history(self.h, level, C.byref(p)))
        return p.value

    def residual(self, r_ptr):
        """r = b - K x of the prepared step at the current state into device memory r_ptr."""
        self.ctx.check(self.L.dhip_ch_residual(self.h, C.c_void_p(r_ptr)))

    def predict_start(self, mlp):
        """(start vector [mu; c] on the host, start kind) for the prepared step."""
        x0 = np.zeros(2 * self.n)
        st = C.c_int32()
        self.ctx.check(self.L.dhip_ch_predict_start(self.h, mlp.h, _vp(x0), HOST, C.byref(st)))
        return x0, st.value

    def solve_learned(self, mlp, raise_on_fail=True):
        """Returns (iterations, residual, converged, start) with start 0 = previous step,
        1 = network prediction, 2 = plain quadratic extrapolation."""
        it, t, st = C.c_int32(), C.c_double(), C.c_int32()
        rc = self.L.dhip_ch_solve_learned(self.h, mlp.h, C.byref(it), C.byref(t), C.byref(st))
        if rc not in (OK, E_NOT_CONVERGED) or (raise_on_fail and rc != OK):
            self.ctx.check(rc)
        return it.value, t.value, rc == OK, st.value

    def step(self, nsteps=1):
        its = np.zeros(nsteps, np.int32)
        res = np.zeros(nsteps)
        self.ctx.check(self.L.dhip_ch_step(self.h, nsteps, _p(its, C.c_int32), _p(res, C.c_double)))
        return its, res

    def set_precision(self, precision):
        """FP64 (default) or FP32_IR: FP32 Arnoldi cycles inside FP64 restarts (dhip_ch_set_precision)."""
        self.ctx.check(self.L.dhip_ch_set_precision(self.h, precision))

    def bdf_step(self, nsteps=1, step_size=0.01, tol=0.005, rho=0.9, min_step=0.0, max_step=float("inf"),
                 adaptive=True, max_tries=0):
        """nsteps of the adaptive BDF1/BDF2 loop (dhip_ch_bdf_step); one dict per step with
        step, t, dt, e, tries, iters_bdf1, iters_bdf2, resid."""
        prm = BdfParams(step_size, tol, rho, min_step, max_step, int(adaptive), max_tries)
        st = (BdfStat * max(nsteps, 1))()
        self.ctx.check(self.L.dhip_ch_bdf_step(self.h, C.byref(prm), nsteps, st))
        keys = ("step", "t", "dt", "e", "tries", "iters_bdf1", "iters_bdf2", "resid")
        return [{k: getattr(st[i], k) for k in keys} for i in range(nsteps)]

    def close(self):
        if self.h and self.ctx.h:
            self.L.dhip_ch_destroy(self.h)
        self.h = None
        self.ctx._csrs.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent P1 problems assembled and solved in one launch (dhip_batch_*).
    meshes: list of KuhnMesh; phi: one float64 array holding every problem's half-grid level set
    (host array, or a device pointer with phi_count and where=DEVICE); offsets: first value of
    each problem in phi; lcs: per problem operator combination (default (1, 1, h) of its mesh)."""

    def __init__(self, ctx, meshes, phi, offsets, coef=(1.0, 1.0, 1.0), rhs_kind=RHS_NONE, rhs_param=None,
                 omit_bound=-1.0, lcs=None, zero_policy=ZERO_THROW, phi_count=None, where=HOST):
        self.ctx, self.L = ctx, ctx.L
        self.n = len(meshes)
        probs = (BatchProblem * self.n)()
        for i, (m, off) in enumerate(zip(meshes, offsets)):
            probs[i].orig[:] = list(m.orig)
            probs[i].e[:] = [float(v) for row in m.e for v in row]
            probs[i].n[:] = list(m.n)
            probs[i].phi_offset = int(off)
            probs[i].lc[:] = list(lcs[i]) if lcs is not None else [1.0, 1.0, m.h]
        co = np.asarray(coef, np.float64)
        rp = None if rhs_param is None else np.asarray(rhs_param, np.float64)
        self.h = C.c_void_p()
        if where == HOST:
            self._phi = np.ascontiguousarray(phi, np.float64)
            ptr, count = _vp(self._phi), self._phi.size
        else:
            ptr, count = C.c_void_p(phi), int(phi_count)
        ctx.check(self.L.dhip_batch_create(ctx.h, self.n, probs, ptr, count, where, zero_policy, _p(co, C.c_double),
                                           rhs_kind, _p(rp, C.c_double), omit_bound, C.byref(self.h)))
        ctx._csrs.add(self)

    def solve(self, max_iter=20000, tol=1e-8, rel=True, graph=False):
        self.ctx.check(self.L.dhip_batch_solve(self.h, int(max_iter), float(tol), int(rel), int(graph)))

    def stats(self):
        """dict of per-problem numpy arrays: status, ncut, ndof, nnz, iters, outer, resid, start_s, end_s"""
        st = (BatchStat * self.n)()
        self.ctx.check(self.L.dhip_batch_stats(self.h, st))
        return {k: np.array([getattr(x, k) for x in st]) for k, _ in BatchStat._fields_}

    def problem(self, p):
        st = self.stats()
        nd, nz = int(st["ndof"][p]), int(st["nnz"][p])
        rp = np.zeros(nd + 1, np.int32)
        col = np.zeros(max(nz, 1), np.int32)
        val = np.zeros(max(nz, 1))
        rhs, x = np.zeros(max(nd, 1)), np.zeros(max(nd, 1))
        self.ctx.check(self.L.dhip_batch_problem(self.h, int(p), _p(rp, C.c_int32), _p(col, C.c_int32),
                                                 _p(val, C.c_double), _p(rhs, C.c_double), _p(x, C.c_double)))
        return dict(rowptr=rp, colind=col[:nz], comb=val[:nz], rhs=rhs[:nd], x=x[:nd])

    def close(self):
        if self.h and self.ctx.h:
            self.L.dhip_batch_destroy(self.h)
        self.h = None
        self.ctx._csrs.discard(self)


class Mlp:
    """The learned warm-start network (dhip_mlp_t); float32 parameters in the layout of
    oracle/mlp_oracle.py, device training and matrix-core inference."""

    def __init__(self, ctx, theta=None):
        self.ctx, self.L = ctx, ctx.L
        self.h = C.c_void_p()
        ctx.check(self.L.dhip_mlp_create(ctx.h, C.byref(self.h)))
        ctx._csrs.add(self)
        if theta is not None:
            self.set_params(theta)

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        assert th.size == MLP_NPARAM
        self.ctx.check(self.L.dhip_mlp_params(self.h, _p(th, C.c_float), None))

    def params(self):
        th = np.zeros(MLP_NPARAM, np.float32)
        self.ctx.check(self.L.dhip_mlp_params(self.h, None, _p(th, C.c_float)))
        return th

    def set_normalization(self, in_mean, in_scale, out_mean, out_scale):
        arr = [np.ascontiguousarray(v, np.float64) for v in (in_mean, in_scale, out_mean, out_scale)]
        self.norm = arr
        self.ctx.check(self.L.dhip_mlp_normalization(self.h, *[_p(v, C.c_double) for v in arr]))

    def features(self, hist_ptrs, rhs_ptr, pcdiag_ptr, n, X_ptr):
        hp = (C.c_void_p * 4)(*[C.c_void_p(p) for p in hist_ptrs] + [None] * (4 - len(hist_ptrs)))
        self.ctx.check(self.L.dhip_mlp_features(self.h, hp, len(hist_ptrs), C.c_void_p(rhs_ptr),
                                                C.c_void_p(pcdiag_ptr), n, C.c_void_p(X_ptr)))

    def targets(self, xnext_ptr, hist_ptrs, n, T_ptr):
        hp = (C.c_void_p * 3)(*[C.c_void_p(p) for p in hist_ptrs] + [None] * (3 - len(hist_ptrs)))
        self.ctx.check(self.L.dhip_mlp_targets(self.h, C.c_void_p(xnext_ptr), hp, len(hist_ptrs), n,
                                               C.c_void_p(T_ptr)))

    def predict(self, X_ptr, S, Y_ptr):
        self.ctx.check(self.L.dhip_mlp_predict(self.h, C.c_void_p(X_ptr), S, C.c_void_p(Y_ptr)))

    def grad(self, X_ptr, T_ptr, idx_ptr, B):
        g = np.zeros(MLP_NPARAM)
        loss = C.c_double()
        self.ctx.check(self.L.dhip_mlp_grad(self.h, C.c_void_p(X_ptr), C.c_void_p(T_ptr), C.c_void_p(idx_ptr), B,
                                            _p(g, C.c_double), C.byref(loss)))
        return loss.value, g

    def train_step(self, X_ptr, T_ptr, idx_ptr, B, lr, want_loss=False):
        loss = C.c_double()
        self.ctx.check(self.L.dhip_mlp_train_step(self.h, C.c_void_p(X_ptr), C.c_void_p(T_ptr), C.c_void_p(idx_ptr),
                                                  B, lr, C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def close(self):
        if self.h and self.ctx.h:
            self.L.dhip_mlp_destroy(self.h)
        self.h = None
        self.ctx._csrs.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    def __init__(self, device=0):
        self.L = load()
        self._csrs = weakref.WeakSet()
        self.h = C.c_void_p()
        rc = self.L.dhip_create(C.byref(self.h), device)
        if rc != OK:
            raise DhipError(rc, "dhip_create failed (no usable GPU?)")

    def check(self, rc):
        if rc != OK:
            raise DhipError(rc, self.L.dhip_last_error(self.h).decode())
        return rc

    def close(self):
        for A in list(self._csrs):
            A.close()
        if self.h:
            self.L.dhip_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- profiling / memory ----
    def profile(self, on):
        self.check(self.L.dhip_profile_enable(self.h, int(on)))

    def profile_get(self, name):
        ms, cnt = C.c_double(), C.c_int64()
        self.check(self.L.dhip_profile_get(self.h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def profile_reset(self):
        self.check(self.L.dhip_profile_reset(self.h))

    def tune(self, key, value):
        self.check(self.L.dhip_tune(self.h, key.encode(), int(value)))

    def solver_probe(self, on):
        """Arm/disarm the solver probe; returns the last solve's ticks as (nblocks, 4) or None."""
        nb = C.c_int32()
        self.check(self.L.dhip_solver_probe(self.h, int(on), None, C.byref(nb)))
        if nb.value <= 0:
            return None
        out = np.zeros((nb.value, 4), np.uint64)
        self.check(self.L.dhip_solver_probe(self.h, int(on), out.ctypes.data_as(C.POINTER(C.c_uint64)), None))
        return out

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.L.dhip_malloc(self.h, C.byref(p), int(nbytes)))
        return p.value

    def free(self, p):
        self.check(self.L.dhip_free(self.h, C.c_void_p(p)))

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.malloc(arr.nbytes)
        self.check(self.L.dhip_memcpy(self.h, C.c_void_p(p), _vp(arr), arr.nbytes, 0))
        return p

    def to_host(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        self.check(self.L.dhip_memcpy(self.h, _vp(out), C.c_void_p(p), out.nbytes, 1))
        return out

    def copy_d2d(self, dst, src, nbytes):
        self.check(self.L.dhip_memcpy(self.h, C.c_void_p(dst), C.c_void_p(src), int(nbytes), 2))

    def synchronize(self):
        self.check(self.L.dhip_synchronize(self.h))

    # ---- mesh / level set ----
    def mesh(self, mesh):
        orig = np.asarray(mesh.orig, np.float64)
        e = np.asarray(mesh.e, np.float64).reshape(9)
        n = np.asarray(mesh.n, np.int32)
        self.check(self.L.dhip_mesh_kuhn(self.h, _p(orig, C.c_double), _p(e, C.c_double), _p(n, C.c_int32)))
        self._mesh = mesh

    def levelset(self, phi=None, device_ptr=None, count=None):
        if device_ptr is not None:
            self.check(self.L.dhip_levelset_p2(self.h, C.c_void_p(device_ptr), int(count), DEVICE))
        else:
            phi = np.ascontiguousarray(phi, np.float64)
            self.check(self.L.dhip_levelset_p2(self.h, _vp(phi), phi.size, HOST))

    # ---- assembly ----
    def classify(self, zero_policy=ZERO_THROW):
        n = C.c_int64()
        self.check(self.L.dhip_classify(self.h, zero_policy, C.byref(n)))
        self.ncut = n.value
        return n.value

    def local_p1(self, coef=(1.0, 1.0, 1.0), rhs_kind=RHS_NONE, rhs_param=None):
        co = np.asarray(coef, np.float64)
        rp = None if rhs_param is None else np.asarray(rhs_param, np.float64)
        self.check(self.L.dhip_local_p1(self.h, _p(co, C.c_double), rhs_kind, _p(rp, C.c_double)))

    def number_p1if(self, omit_bound=-1.0):
        n = C.c_int64()
        self.check(self.L.dhip_number_p1if(self.h, omit_bound, C.byref(n)))
        self.ndof = n.value
        return n.value

    def build_p1(self, lc=(1.0, 1.0, 1.0)):
        lcv = np.asarray(lc, np.float64)
        n = C.c_int64()
        self.check(self.L.dhip_build_p1(self.h, _p(lcv, C.c_double), C.byref(n)))
        self.nnz = n.value
        return n.value

    def assemble_p1(self, zero_policy=ZERO_THROW, coef=(1.0, 1.0, 1.0), rhs_kind=RHS_NONE, rhs_param=None,
                    omit_bound=-1.0, lc=(1.0, 1.0, 1.0)):
        co = np.asarray(coef, np.float64)
        rp = None if rhs_param is None else np.asarray(rhs_param, np.float64)
        lcv = np.asarray(lc, np.float64)
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self.check(self.L.dhip_assemble_p1(self.h, zero_policy, _p(co, C.c_double), rhs_kind, _p(rp, C.c_double),
                                           omit_bound, _p(lcv, C.c_double), C.byref(a), C.byref(b), C.byref(c)))
        self.ncut, self.ndof, self.nnz, self.space = a.value, b.value, c.value, 1
        return a.value, b.value, c.value

    def assemble_p1_band(self, width, zero_policy=ZERO_THROW, coef=(1.0, 1.0, 1.0), rhs_kind=RHS_NONE,
                         rhs_param=None, lc=(1.0, 1.0, 1.0)):
        """Narrow-band P1 assembly: (nband, ndof, nnz); results through csr(), cut_tets() (the band
        tets), local() (flags bit 5: band tet without patch), dof_vertices()."""
        co = np.asarray(coef, np.float64)
        rp = None if rhs_param is None else np.asarray(rhs_param, np.float64)
        lcv = np.asarray(lc, np.float64)
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self.check(self.L.dhip_assemble_p1_band(self.h, zero_policy, _p(co, C.c_double), rhs_kind,
                                                _p(rp, C.c_double), float(width), _p(lcv, C.c_double), C.byref(a),
                                                C.byref(b), C.byref(c)))
        self.ncut, self.ndof, self.nnz, self.space = a.value, b.value, c.value, 1
        return a.value, b.value, c.value

    def recovered_gradient(self):
        """averaging_P2_gradient_recovery of the current level set: shape (half-grid nodes, 3)"""
        n = self._mesh.n
        nn = (2 * n[0] + 1) * (2 * n[1] + 1) * (2 * n[2] + 1)
        g = np.zeros(3 * nn)
        self.check(self.L.dhip_p2_recovered_gradient(self.h, _p(g, C.c_double), HOST))
        return g.reshape(nn, 3)

    def quaqua_points(self, params=None):
        """base points (npoints, 3) of every mapped quadrature point of the cut tets (after classify())"""
        P = params if params is not None else QuaQuaParams()
        n = C.c_int64()
        self.check(self.L.dhip_quaqua_points(self.h, C.byref(P), C.byref(n), None, HOST))
        xyz = np.zeros((max(n.value, 1), 3))
        self.check(self.L.dhip_quaqua_points(self.h, C.byref(P), C.byref(n), _p(xyz, C.c_double), HOST))
        return xyz[:n.value]

    def quaqua_values(self, values):
        v = np.ascontiguousarray(values, np.float64)
        self.check(self.L.dhip_quaqua_values(self.h, _p(v, C.c_double), v.size, HOST))

    def assemble_p2_quaqua(self, zero_policy=ZERO_THROW, params=None, rhs_kind=RHS_NONE, rhs_param=None,
                           omit_bound=-1.0, lc=(1.0, 1.0, 0.0)):
        """P2 mass (slot 0), Laplace-Beltrami (slot 1) and load on the QuaQua-mapped surface, their combination
        lc[0] M + lc[1] A: (ncut, ndof, nnz); results through csr()."""
        rp = None if rhs_param is None else np.asarray(rhs_param, np.float64)
        lcv = np.asarray(lc, np.float64)
        P = params if params is not None else QuaQuaParams()
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self.check(self.L.dhip_assemble_p2_quaqua(self.h, zero_policy, C.byref(P), rhs_kind, _p(rp, C.c_double),
                                                  omit_bound, _p(lcv, C.c_double), C.byref(a), C.byref(b),
                                                  C.byref(c)))
        self.ncut, self.ndof, self.nnz, self.space = a.value, b.value, c.value, 2
        return self.ncut, self.ndof, self.nnz

    def assemble_p2(self, zero_policy=ZERO_THROW, coef=(1.0, 1.0, 1.0), rhs_kind=RHS_NONE, rhs_param=None,
                    omit_bound=-1.0, lc=(1.0, 1.0, 1.0)):
        """P2 interface space: (ncut, ndof, nnz); results through csr(), tet_dofs(), dof_nodes()."""
        co = np.asarray(coef, np.float64)
        rp = None if rhs_param is None else np.asarray(rhs_param, np.float64)
        lcv = np.asarray(lc, np.float64)
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self.check(self.L.dhip_assemble_p2(self.h, zero_policy, _p(co, C.c_double), rhs_kind, _p(rp, C.c_double),
                                           omit_bound, _p(lcv, C.c_double), C.byref(a), C.byref(b), C.byref(c)))
        self.ncut, self.ndof, self.nnz, self.space = a.value, b.value, c.value, 2
        return a.value, b.value, c.value

    def local_p2(self):
        loc = np.zeros((max(self.ncut, 1), LOC2_N))
        fl = np.zeros(max(self.ncut, 1), np.uint8)
        self.check(self.L.dhip_get_local_p2(self.h, _p(loc, C.c_double), _p(fl, C.c_uint8)))
        return loc[: self.ncut], fl[: self.ncut]

    def dof_nodes(self):
        out = np.zeros(max(self.ndof, 1), np.int64)
        self.check(self.L.dhip_get_dof_nodes(self.h, _p(out, C.c_int64)))
        return out[: self.ndof]

    def cut_tets(self):
        out = np.zeros(max(self.ncut, 1), np.int64)
        self.check(self.L.dhip_get_cut_tets(self.h, _p(out, C.c_int64)))
        return out[: self.ncut]

    def local(self):
        loc = np.zeros((max(self.ncut, 1), LOC_N))
        fl = np.zeros(max(self.ncut, 1), np.uint8)
        self.check(self.L.dhip_get_local(self.h, _p(loc, C.c_double), _p(fl, C.c_uint8)))
        return loc[: self.ncut], fl[: self.ncut]

    def tet_dofs(self):
        out = np.zeros((max(self.ncut, 1), 10 if getattr(self, "space", 1) == 2 else 4), np.int32)
        self.check(self.L.dhip_get_tet_dofs(self.h, _p(out, C.c_int32)))
        return out[: self.ncut]

    def dof_vertices(self):
        out = np.zeros(max(self.ndof, 1), np.int64)
        self.check(self.L.dhip_get_dof_vertices(self.h, _p(out, C.c_int64)))
        return out[: self.ndof]

    def csr(self):
        n, z = self.ndof, self.nnz
        rp = np.zeros(n + 1, np.int32)
        col = np.zeros(max(z, 1), np.int32)
        vals = [np.zeros(max(z, 1)) for _ in range(4)]
        rhs = np.zeros(max(n, 1))
        self.check(self.L.dhip_get_csr(self.h, _p(rp, C.c_int32), _p(col, C.c_int32),
                                       *[_p(v, C.c_double) for v in vals], _p(rhs, C.c_double)))
        return dict(rowptr=rp, colind=col[:z], mass=vals[0][:z], lb=vals[1][:z], stab=vals[2][:z],
                    comb=vals[3][:z], rhs=rhs[:n])

    def device_arrays(self):
        ps = [C.c_void_p() for _ in range(4)]
        self.check(self.L.dhip_device_arrays(self.h, *[C.byref(p) for p in ps]))
        return [p.value for p in ps]

    # ---- matrices / vectors ----
    def csr_create(self, rowptr, colind, val):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        colind = np.ascontiguousarray(colind, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        h = C.c_void_p()
        self.check(self.L.dhip_csr_create(self.h, rowptr.size - 1, colind.size, _vp(rowptr), _vp(colind), _vp(val),
                                          HOST, C.byref(h)))
        return Csr(self, h)

    def csr_import(self, path, fmt="mtx"):
        """a square matrix file (MatrixMarket or plain) into a new device matrix"""
        h = C.c_void_p()
        self.check(self.L.dhip_csr_import(self.h, os.fsencode(path), _fmt(fmt), C.byref(h)))
        return Csr(self, h)

    def vec_write(self, path, x_ptr, n, binary=True):
        """VecDescBaseCL::Write of a device vector (x_ptr: device address)"""
        self.check(self.L.dhip_vec_write(self.h, os.fsencode(path), _vp(x_ptr), int(n), DEVICE, int(binary)))

    def vec_read(self, path, x_ptr, n, binary=True):
        """VecDescBaseCL::Read into a device vector of n values"""
        self.check(self.L.dhip_vec_read(self.h, os.fsencode(path), _vp(x_ptr), int(n), DEVICE, int(binary)))

    def csr_from_assembly(self, op=OP_COMB):
        h = C.c_void_p()
        self.check(self.L.dhip_csr_from_assembly(self.h, op, C.byref(h)))
        return Csr(self, h)

    def spmv(self, A, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(A.shape[0])
        self.check(self.L.dhip_spmv(self.h, A.h, _vp(x), _vp(y), HOST))
        return y

    def quad_points(self):
        """(npoints, 3) world coordinates of the patches' composite Quad5 points (dhip_quad_points)."""
        n = C.c_int64()
        self.check(self.L.dhip_quad_points(self.h, C.byref(n), None, HOST))
        xyz = np.zeros((max(n.value, 1), 3))
        if n.value:
            self.check(self.L.dhip_quad_points(self.h, C.byref(n), _vp(xyz), HOST))
        return xyz[:n.value]

    def rhs_values(self, values):
        """f at the points of quad_points() for the next DHIP_RHS_VALUES assembly (dhip_rhs_values)."""
        v = np.ascontiguousarray(values, np.float64)
        self.check(self.L.dhip_rhs_values(self.h, _vp(v) if v.size else None, v.size, HOST))

    def spmv_t(self, A, x, ncols=None):
        """A^T x (dhip_spmv_t); ncols defaults to the row count (square A)."""
        x = np.ascontiguousarray(x, np.float64)
        m = A.shape[0] if ncols is None else ncols
        y = np.zeros(m)
        self.check(self.L.dhip_spmv_t(self.h, A.h, m, _vp(x), _vp(y), HOST))
        return y

    def dot(self, x, y):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        r = C.c_double()
        self.check(self.L.dhip_dot(self.h, x.size, _vp(x), _vp(y), C.byref(r), HOST))
        return r.value

    def axpy(self, a, x, y):
        x = np.ascontiguousarray(x, np.float64)
        y = np.array(y, np.float64)
        self.check(self.L.dhip_axpy(self.h, x.size, a, _vp(x), _vp(y), HOST))
        return y

    def pcg(self, A, b, x0=None, pc=PC_JACOBI, max_iter=10000, tol=1e-10, rel=True, precision=FP64,
            where=HOST, raise_on_fail=True):
        """Returns (x, iterations, residual, converged).  With where=DEVICE, b and x0 are
        device pointers and x0 is overwritten with the solution."""
        it = C.c_int32(max_iter)
        t = C.c_double(tol)
        if where == DEVICE:
            rc = self.L.dhip_pcg(self.h, A.h, pc, C.c_void_p(b), C.c_void_p(x0), DEVICE, C.byref(it), C.byref(t),
                                 int(rel), precision)
            x = None
        else:
            b = np.ascontiguousarray(b, np.float64)
            x = np.zeros(b.size) if x0 is None else np.array(x0, np.float64)
            rc = self.L.dhip_pcg(self.h, A.h, pc, _vp(b), _vp(x), HOST, C.byref(it), C.byref(t), int(rel), precision)
        if rc not in (OK, E_NOT_CONVERGED) or (raise_on_fail and rc != OK):
            self.check(rc)
        return x, it.value, t.value, rc == OK

    def gmres_diag(self, A, pcdiag, b, x0=None, restart=50, max_iter=10000, tol=1e-10, rel=True, left_pc=False,
                   raise_on_fail=True):
        """GMRES preconditioned by z = r / pcdiag (block Jacobi with a given diagonal)."""
        b = np.ascontiguousarray(b, np.float64)
        d = np.ascontiguousarray(pcdiag, np.float64)
        x = np.zeros(b.size) if x0 is None else np.array(x0, np.float64)
        it = C.c_int32(max_iter)
        t = C.c_double(tol)
        rc = self.L.dhip_gmres_diag(self.h, A.h, _vp(d), _vp(b), _vp(x), HOST, restart, C.byref(it), C.byref(t),
                                    int(rel), int(left_pc))
        if rc not in (OK, E_NOT_CONVERGED) or (raise_on_fail and rc != OK):
            self.check(rc)
        return x, it.value, t.value, rc == OK

    def cahn_hilliard(self, dt, epsilon, rho, **kw):
        return CahnHilliard(self, dt, epsilon, rho, **kw)

    def batch(self, meshes, phi, offsets, **kw):
        return Batch(self, meshes, phi, offsets, **kw)

    def mlp(self, theta=None):
        return Mlp(self, theta)

    def gmres(self, A, b, x0=None, pc=PC_JACOBI, restart=50, max_iter=10000, tol=1e-10, rel=True, left_pc=False,
              raise_on_fail=True):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(b.size) if x0 is None else np.array(x0, np.float64)
        it = C.c_int32(max_iter)
        t = C.c_double(tol)
        rc = self.L.dhip_gmres(self.h, A.h, pc, _vp(b), _vp(x), HOST, restart, C.byref(it), C.byref(t), int(rel),
                               int(left_pc))
        if rc not in (OK, E_NOT_CONVERGED) or (raise_on_fail and rc != OK):
            self.check(rc)
        return x, it.value, t.value, rc == OK
