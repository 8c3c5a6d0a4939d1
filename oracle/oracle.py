"""ctypes wrapper of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, to check results against.  It loads oracle/_build/liboracle.so (parity build,
no FMA contraction) or liboracle_omp.so (OpenMP build with the reference's
release flags, used for CPU timing).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

OK, E_ZERO_NODE, E_ZERO_PATTERN, E_ARG, E_NOT_CONVERGED, E_NOMEM, E_CAPACITY = range(7)
ZERO_THROW, ZERO_ALLOW = 0, 1
RHS_NONE, RHS_3XY_MINUS_X3, RHS_CONST = 0, 1, 2


class OracleError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"oracle {what} failed with code {code}")
        self.code = code


class Mesh(C.Structure):
    _fields_ = [("orig", C.c_double * 3), ("e", (C.c_double * 3) * 3), ("n", C.c_int32 * 3)]


class Local(C.Structure):
    _fields_ = [("nfacet", C.c_int32), ("nvert", C.c_int32), ("area", C.c_double),
                ("mass", C.c_double * 16), ("lb", C.c_double * 16), ("stab", C.c_double * 16),
                ("rhs", C.c_double * 4), ("hat_sq", C.c_double * 4)]


class LocalP2(C.Structure):
    _fields_ = [("nfacet", C.c_int32), ("mass", C.c_double * 100), ("lb", C.c_double * 100),
                ("stab", C.c_double * 100), ("rhs", C.c_double * 10), ("hat_sq", C.c_double * 10)]


class Patch(C.Structure):
    _fields_ = [("nfacet", C.c_int32), ("bnd", C.c_int32 * 16), ("world", ((C.c_double * 3) * 3) * 16),
                ("bary", ((C.c_double * 4) * 3) * 16), ("absdet", C.c_double * 16),
                ("normal", (C.c_double * 3) * 16), ("ls", C.c_double * 10)]


class BdfParams(C.Structure):
    """orc_bdf_params_t (surfphasesep/surfphasesep.cpp:112-123)"""
    _fields_ = [("step_size", C.c_double), ("tol", C.c_double), ("rho", C.c_double), ("min_step", C.c_double),
                ("max_step", C.c_double), ("adaptive", C.c_int32), ("max_tries", C.c_int32)]


class BdfState(C.Structure):
    """orc_bdf_state_t"""
    _fields_ = [("step", C.c_int64), ("t", C.c_double), ("dt", C.c_double), ("e", C.c_double)]


class BdfStat(C.Structure):
    """orc_bdf_stat_t"""
    _fields_ = [("tries", C.c_int32), ("iters_bdf1", C.c_int32), ("iters_bdf2", C.c_int32), ("resid", C.c_double),
                ("dt", C.c_double), ("e", C.c_double)]


class QuaQua(C.Structure):
    """orc_quaqua_t: the QuaQuaMapperCL parameters (LevelsetMapper.Iter / Tol / ArmijoConstant, damping steps)"""
    _fields_ = [("maxiter", C.c_int32), ("tol", C.c_double), ("armijo_c", C.c_double), ("max_damping", C.c_int32)]


def quaqua_params(maxiter=20, tol=1e-7, armijo_c=1e-2, max_damping=8):
    """the reference's parameter files (param/surfactant/surfactant/surfactant.json: LevelsetMapper)"""
    return QuaQua(maxiter, tol, armijo_c, max_damping)


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


_libs = {}


def lib(kind="parity"):
    name = "liboracle.so" if kind == "parity" else "liboracle_omp.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(HERE, "_build", name)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
    L = C.CDLL(path)
    D, I32, I64 = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    M = C.POINTER(Mesh)
    L.orc_classify.argtypes = [M, D, C.c_int, C.c_int64, I64, I64]
    L.orc_number_p1if.argtypes = [M, D, C.c_double, C.c_int, I32, I64]
    L.orc_tet_local.argtypes = [M, D, C.c_int64, D, C.c_int, D, C.c_int, C.POINTER(Local)]
    L.orc_assemble_p1.argtypes = [M, D, I32, C.c_int64, D, C.c_int, D, C.c_int, C.c_int64, I64,
                                  I64, I32, D, D, D, D]
    L.orc_number_p2if.argtypes = [M, D, C.c_double, C.c_int, I32, I64]
    L.orc_tet_local_p2.argtypes = [M, D, C.c_int64, D, C.c_int, D, C.c_int, C.POINTER(LocalP2)]
    L.orc_assemble_p2.argtypes = [M, D, I32, C.c_int64, D, C.c_int, D, C.c_int, C.c_int64, I64,
                                  I64, I32, D, D, D, D]
    L.orc_surface_area.argtypes = [M, D, C.c_int, D]
    L.orc_color_classes.argtypes = [M, I64, I64, I64]
    L.orc_assemble_p1_colored.argtypes = [M, D, I32, C.c_int64, D, C.c_int, D, C.c_int, C.c_int64, I64, I64,
                                          C.c_int64, I64, I64, I32, D, D, D, D]
    L.orc_tet_patch.argtypes = [M, D, C.c_int64, C.c_int, C.POINTER(Patch)]
    L.orc_spmv.argtypes = [C.c_int64, I64, I32, D, D, D]
    L.orc_kahan_dot.argtypes = [C.c_int64, D, D]
    L.orc_kahan_dot.restype = C.c_double
    L.orc_pcg_jacobi.argtypes = [C.c_int64, I64, I32, D, D, D, I32, D, C.c_int32]
    L.orc_gmres_jacobi.argtypes = [C.c_int64, I64, I32, D, D, D, C.c_int32, I32, D, C.c_int32,
                                   C.c_int32, C.c_int32]
    L.orc_gmres_diag.argtypes = [C.c_int64, I64, I32, D, D, D, D, C.c_int32, I32, D, C.c_int32, C.c_int32]
    L.orc_tet_mobility.argtypes = [M, D, C.c_int64, D, C.c_int, D]
    L.orc_tet_mobility_p2.argtypes = [M, D, C.c_int64, D, C.c_int, D]
    L.orc_ch_assemble.argtypes = [M, D, I32, C.c_int64, I64, I32, D, C.c_int, D, D, D]
    L.orc_ch_assemble_p2.argtypes = [M, D, I32, C.c_int64, I64, I32, D, C.c_int, D, D, D]
    L.orc_ch_system.argtypes = [M, D, I32, C.c_int64, I64, I32, D, D, D, D, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_double, I64, I32, D, D, D]
    L.orc_ch_system_p2.argtypes = [M, D, I32, C.c_int64, I64, I32, D, D, D, D, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_double, I64, I32, D, D, D]
    L.orc_ch_step.argtypes = [M, D, I32, C.c_int64, I64, I32, D, D, D, D, C.c_double, C.c_double, C.c_double,
                              C.c_double, C.c_double, C.c_int32, I32, D, C.c_int32]
    L.orc_ch_step_p2.argtypes = [M, D, I32, C.c_int64, I64, I32, D, D, D, D, C.c_double, C.c_double, C.c_double,
                              C.c_double, C.c_double, C.c_int32, I32, D, C.c_int32]
    BP, BS, BT = C.POINTER(BdfParams), C.POINTER(BdfState), C.POINTER(BdfStat)
    L.orc_bdf2_weights.argtypes = [C.c_double, C.c_double, D]
    L.orc_bdf2_weights.restype = None
    L.orc_bdf_step_size.argtypes = [BP, C.c_double, C.c_double]
    L.orc_bdf_step_size.restype = C.c_double
    L.orc_ch_system_bdf.argtypes = [C.c_int, M, D, I32, C.c_int64, I64, I32, D, D, D, D, D, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_double, I64, I32, D, D, D]
    L.orc_ch_bdf_step.argtypes = [C.c_int, M, D, I32, C.c_int64, I64, I32, D, D, D, D, D, BS, BP, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_double,
                                  C.c_int32, BT]
    L.orc_lattice_tables.argtypes = [I32, I32]
    L.orc_quad_tables.argtypes = [D, D, D, D]
    L.orc_classify_band.argtypes = [M, D, C.c_double, C.c_int, C.c_int64, I64, I64, C.POINTER(C.c_uint8)]
    L.orc_number_p1band.argtypes = [M, D, C.c_double, C.c_int, I32, I64]
    L.orc_assemble_p1band.argtypes = [M, D, I32, C.c_int64, C.c_double, D, C.c_int, D, C.c_int, C.c_int64,
                                      I64, I64, I32, D, I64, I64, I32, D, D, D]
    L.orc_refpatch_table.argtypes = [I32, I32, I32]
    QQ = C.POINTER(QuaQua)
    L.orc_p2_recovered_gradient.argtypes = [M, D, D]
    L.orc_quaqua_point.argtypes = [M, D, D, C.c_int64, D, QQ, D, D, D, D, I32]
    L.orc_quaqua_winv.argtypes = [M, D, D, C.c_int64, D, D, QQ, D]
    L.orc_tet_local_p2qq.argtypes = [M, D, D, C.c_int64, QQ, C.c_int, D, C.c_int, C.POINTER(LocalP2)]
    L.orc_assemble_p2qq.argtypes = [M, D, D, I32, C.c_int64, QQ, C.c_int, D, C.c_int, C.c_int64, I64, I64, I32,
                                    D, D, D]
    _libs[name] = L
    return L


def make_mesh(orig, e, n):
    m = Mesh()
    for c in range(3):
        m.orig[c] = float(orig[c])
        m.n[c] = int(n[c])
        for k in range(3):
            m.e[k][c] = float(e[k][c])
    return m


def _check(code, what):
    if code != OK:
        raise OracleError(code, what)


def lattice_tables():
    b = np.zeros((10, 4), np.int32)
    t = np.zeros((8, 4), np.int32)
    lib().orc_lattice_tables(_ptr(b, C.c_int32), _ptr(t, C.c_int32))
    return b, t


def quad_tables():
    """(q5n (7,3), q5w (7,), v5n (15,4), (C1, C2)) of the oracle's quadrature rules"""
    q5n, q5w, v5n, v5c = np.zeros((7, 3)), np.zeros(7), np.zeros((15, 4)), np.zeros(2)
    lib().orc_quad_tables(*[_ptr(a, C.c_double) for a in (q5n, q5w, v5n, v5c)])
    return q5n, q5w, v5n, tuple(v5c)


def refpatch_table():
    size = np.zeros(81, np.int32)
    tri = np.zeros((81, 2, 3), np.int32)
    bnd = np.zeros(81, np.int32)
    lib().orc_refpatch_table(_ptr(size, C.c_int32), _ptr(tri, C.c_int32), _ptr(bnd, C.c_int32))
    return size, tri, bnd


def classify(mesh, phi, zero_policy=ZERO_THROW):
    phi = np.ascontiguousarray(phi, np.float64)
    n = C.c_int64()
    _check(lib().orc_classify(C.byref(mesh), _ptr(phi, C.c_double), zero_policy, 0, C.byref(n), None),
           "classify")
    out = np.zeros(max(n.value, 1), np.int64)
    _check(lib().orc_classify(C.byref(mesh), _ptr(phi, C.c_double), zero_policy, n.value, C.byref(n),
                              _ptr(out, C.c_int64)), "classify")
    return out[: n.value]


def number_p1if(mesh, phi, omit_bound=-1.0, zero_policy=ZERO_THROW, kind="parity"):
    phi = np.ascontiguousarray(phi, np.float64)
    nv = (mesh.n[0] + 1) * (mesh.n[1] + 1) * (mesh.n[2] + 1)
    dof = np.empty(nv, np.int32)
    nd = C.c_int64()
    _check(lib(kind).orc_number_p1if(C.byref(mesh), _ptr(phi, C.c_double), omit_bound, zero_policy,
                                 _ptr(dof, C.c_int32), C.byref(nd)), "number_p1if")
    return dof, nd.value


def tet_local(mesh, phi, tet, coef=(1.0, 1.0, 1.0), rhs_kind=RHS_NONE, rhs_param=None,
              zero_policy=ZERO_THROW):
    phi = np.ascontiguousarray(phi, np.float64)
    co = np.asarray(coef, np.float64)
    rp = None if rhs_param is None else np.asarray(rhs_param, np.float64)
    out = Local()
    _check(lib().orc_tet_local(C.byref(mesh), _ptr(phi, C.c_double), int(tet), _ptr(co, C.c_double),
                               rhs_kind, _ptr(rp, C.c_double), zero_policy, C.byref(out)), "tet_local")
    return out


def tet_patch(mesh, phi, tet, zero_policy=ZERO_THROW):
    """dict(nfacet, bnd, world (F,3,3), bary (F,3,4), absdet (F,), normal (F,3), ls (10,))"""
    phi = np.ascontiguousarray(phi, np.float64)
    p = Patch()
    _check(lib().orc_tet_patch(C.byref(mesh), _ptr(phi, C.c_double), int(tet), zero_policy, C.byref(p)),
           "tet_patch")
    f = p.nfacet
    return dict(nfacet=f, bnd=np.ctypeslib.as_array(p.bnd)[:f].copy(),
                world=np.ctypeslib.as_array(p.world)[:f].copy(), bary=np.ctypeslib.as_array(p.bary)[:f].copy(),
                absdet=np.ctypeslib.as_array(p.absdet)[:f].copy(), normal=np.ctypeslib.as_array(p.normal)[:f].copy(),
                ls=np.ctypeslib.as_array(p.ls).copy())


def assemble_p1(mesh, phi, dof_of_vertex, ndof, coef=(1.0, 1.0, 1.0), rhs_kind=RHS_NONE,
                rhs_param=None, zero_policy=ZERO_THROW, kind="parity"):
    """Returns dict(rowptr, colind, mass, lb, stab, rhs) -- CSR with int64 row pointers."""
    phi = np.ascontiguousarray(phi, np.float64)
    co = np.asarray(coef, np.float64)
    rp = None if rhs_param is None else np.asarray(rhs_param, np.float64)
    L = lib(kind)
    ncut = classify(mesh, phi, zero_policy).size
    cap = 16 * max(ncut, 1)
    rowptr = np.zeros(ndof + 1, np.int64)
    col = np.zeros(cap, np.int32)
    vm, vl, vs = (np.zeros(cap) for _ in range(3))
    rhs = np.zeros(max(ndof, 1))
    nnz = C.c_int64()
    _check(L.orc_assemble_p1(C.byref(mesh), _ptr(phi, C.c_double), _ptr(dof_of_vertex, C.c_int32), ndof,
                             _ptr(co, C.c_double), rhs_kind, _ptr(rp, C.c_double), zero_policy, cap,
                             C.byref(nnz), _ptr(rowptr, C.c_int64), _ptr(col, C.c_int32),
                             _ptr(vm, C.c_double), _ptr(vl, C.c_double), _ptr(vs, C.c_double),
                             _ptr(rhs, C.c_double)), "assemble_p1")
    k = nnz.value
    return dict(rowptr=rowptr, colind=col[:k].copy(), mass=vm[:k].copy(), lb=vl[:k].copy(),
                stab=vs[:k].copy(), rhs=rhs[:ndof].copy())


def color_classes(mesh, kind="parity"):
    """ColorClassesCL of the brick: (class_start[ncolors + 1], class_tets[6 n^3])"""
    nt = 6 * mesh.n[0] * mesh.n[1] * mesh.n[2]
    start = np.zeros(129, np.int64)
    tets = np.zeros(max(nt, 1), np.int64)
    nc = C.c_int64()
    _check(lib(kind).orc_color_classes(C.byref(mesh), C.byref(nc), _ptr(start, C.c_int64), _ptr(tets, C.c_int64)),
           "color_classes")
    return start[:nc.value + 1].copy(), tets[:nt]


def assemble_p1_colored(mesh, phi, dof_of_vertex, ndof, colors, coef=(1.0, 1.0, 1.0), rhs_kind=RHS_NONE,
                        rhs_param=None, zero_policy=ZERO_THROW, kind="omp", cap=None):
    """The colour-parallel accumulate() (OpenMP build by default): the same matrices as
    assemble_p1 up to the order in which each entry's contributions are summed."""
    phi = np.ascontiguousarray(phi, np.float64)
    co = np.asarray(coef, np.float64)
    rp = None if rhs_param is None else np.asarray(rhs_param, np.float64)
    start, tets = colors
    if cap is None:
        cap = 16 * max(ndof, 1)
    rowptr = np.zeros(ndof + 1, np.int64)
    col = np.zeros(cap, np.int32)
    vm, vl, vs = (np.zeros(cap) for _ in range(3))
    rhs = np.zeros(max(ndof, 1))
    nnz = C.c_int64()
    _check(lib(kind).orc_assemble_p1_colored(C.byref(mesh), _ptr(phi, C.c_double), _ptr(dof_of_vertex, C.c_int32),
                                             ndof, _ptr(co, C.c_double), rhs_kind, _ptr(rp, C.c_double), zero_policy,
                                             start.size - 1, _ptr(start, C.c_int64), _ptr(tets, C.c_int64), cap,
                                             C.byref(nnz), _ptr(rowptr, C.c_int64), _ptr(col, C.c_int32),
                                             _ptr(vm, C.c_double), _ptr(vl, C.c_double), _ptr(vs, C.c_double),
                                             _ptr(rhs, C.c_double)), "assemble_p1_colored")
    k = nnz.value
    return dict(rowptr=rowptr, colind=col[:k].copy(), mass=vm[:k].copy(), lb=vl[:k].copy(),
                stab=vs[:k].copy(), rhs=rhs[:ndof].copy())


def number_p2if(mesh, phi, omit_bound=-1.0, zero_policy=ZERO_THROW, kind="parity"):
    """dof number of every half-grid node (-1: not an interface DOF) and the DOF count"""
    phi = np.ascontiguousarray(phi, np.float64)
    nn = (2 * mesh.n[0] + 1) * (2 * mesh.n[1] + 1) * (2 * mesh.n[2] + 1)
    dof = np.empty(nn, np.int32)
    nd = C.c_int64()
    _check(lib(kind).orc_number_p2if(C.byref(mesh), _ptr(phi, C.c_double), omit_bound, zero_policy,
                                     _ptr(dof, C.c_int32), C.byref(nd)), "number_p2if")
    return dof, nd.value


def tet_local_p2(mesh, phi, tet, coef=(1.0, 1.0, 1.0), rhs_kind=RHS_NONE, rhs_param=None,
                 zero_policy=ZERO_THROW):
    phi = np.ascontiguousarray(phi, np.float64)
    co = np.asarray(coef, np.float64)
    rp = None if rhs_param is None else np.asarray(rhs_param, np.float64)
    out = LocalP2()
    _check(lib().orc_tet_local_p2(C.byref(mesh), _ptr(phi, C.c_double), int(tet), _ptr(co, C.c_double),
                                  rhs_kind, _ptr(rp, C.c_double), zero_policy, C.byref(out)), "tet_local_p2")
    return out


def assemble_p2(mesh, phi, dof_of_node, ndof, coef=(1.0, 1.0, 1.0), rhs_kind=RHS_NONE,
                rhs_param=None, zero_policy=ZERO_THROW, kind="parity"):
    """P2 interface matrices: dict(rowptr, colind, mass, lb, stab, rhs), int64 row pointers."""
    phi = np.ascontiguousarray(phi, np.float64)
    co = np.asarray(coef, np.float64)
    rp = None if rhs_param is None else np.asarray(rhs_param, np.float64)
    L = lib(kind)
    ncut = classify(mesh, phi, zero_policy).size
    cap = 100 * max(ncut, 1)
    rowptr = np.zeros(ndof + 1, np.int64)
    col = np.zeros(cap, np.int32)
    vm, vl, vs = (np.zeros(cap) for _ in range(3))
    rhs = np.zeros(max(ndof, 1))
    nnz = C.c_int64()
    _check(L.orc_assemble_p2(C.byref(mesh), _ptr(phi, C.c_double), _ptr(dof_of_node, C.c_int32), ndof,
                             _ptr(co, C.c_double), rhs_kind, _ptr(rp, C.c_double), zero_policy, cap,
                             C.byref(nnz), _ptr(rowptr, C.c_int64), _ptr(col, C.c_int32),
                             _ptr(vm, C.c_double), _ptr(vl, C.c_double), _ptr(vs, C.c_double),
                             _ptr(rhs, C.c_double)), "assemble_p2")
    k = nnz.value
    return dict(rowptr=rowptr, colind=col[:k].copy(), mass=vm[:k].copy(), lb=vl[:k].copy(),
                stab=vs[:k].copy(), rhs=rhs[:ndof].copy())


def recovered_gradient(mesh, phi):
    """averaging_P2_gradient_recovery: the recovered gradient at every half-grid node, shape (nodes, 3)"""
    phi = np.ascontiguousarray(phi, np.float64)
    nn = (2 * mesh.n[0] + 1) * (2 * mesh.n[1] + 1) * (2 * mesh.n[2] + 1)
    g = np.zeros(3 * nn)
    _check(lib().orc_p2_recovered_gradient(C.byref(mesh), _ptr(phi, C.c_double), _ptr(g, C.c_double)),
           "p2_recovered_gradient")
    return g.reshape(nn, 3)


def quaqua_point(mesh, phi, grad, tet, xb, params=None):
    """base point (barycentric in the tet, world), dh, Jacobian dph (3x3) and Newton iterations"""
    phi = np.ascontiguousarray(phi, np.float64)
    g = np.ascontiguousarray(grad, np.float64).reshape(-1)
    xb = np.ascontiguousarray(xb, np.float64)
    P = params if params is not None else quaqua_params()
    bxb, y, dh, dph, it = np.zeros(4), np.zeros(3), C.c_double(), np.zeros(9), C.c_int32()
    _check(lib().orc_quaqua_point(C.byref(mesh), _ptr(phi, C.c_double), _ptr(g, C.c_double), int(tet),
                                  _ptr(xb, C.c_double), C.byref(P), _ptr(bxb, C.c_double), _ptr(y, C.c_double),
                                  C.byref(dh), _ptr(dph, C.c_double), C.byref(it)), "quaqua_point")
    return bxb, y, dh.value, dph.reshape(3, 3), it.value


def quaqua_winv(mesh, phi, grad, tet, xb, nl, params=None):
    """W^-1 (3x3) of the gradient transformation at xb for the plane normal nl"""
    phi = np.ascontiguousarray(phi, np.float64)
    g = np.ascontiguousarray(grad, np.float64).reshape(-1)
    xb = np.ascontiguousarray(xb, np.float64)
    nl = np.ascontiguousarray(nl, np.float64)
    P = params if params is not None else quaqua_params()
    w = np.zeros(9)
    _check(lib().orc_quaqua_winv(C.byref(mesh), _ptr(phi, C.c_double), _ptr(g, C.c_double), int(tet),
                                 _ptr(xb, C.c_double), _ptr(nl, C.c_double), C.byref(P), _ptr(w, C.c_double)),
           "quaqua_winv")
    return w.reshape(3, 3)


def tet_local_p2qq(mesh, phi, grad, tet, params=None, rhs_kind=RHS_NONE, rhs_param=None, zero_policy=ZERO_THROW):
    phi = np.ascontiguousarray(phi, np.float64)
    g = np.ascontiguousarray(grad, np.float64).reshape(-1)
    rp = None if rhs_param is None else np.asarray(rhs_param, np.float64)
    P = params if params is not None else quaqua_params()
    out = LocalP2()
    _check(lib().orc_tet_local_p2qq(C.byref(mesh), _ptr(phi, C.c_double), _ptr(g, C.c_double), int(tet),
                                    C.byref(P), rhs_kind, _ptr(rp, C.c_double), zero_policy, C.byref(out)),
           "tet_local_p2qq")
    return out


def assemble_p2qq(mesh, phi, grad, dof_of_node, ndof, params=None, rhs_kind=RHS_NONE, rhs_param=None,
                  zero_policy=ZERO_THROW, kind="parity"):
    """P2 mass, Laplace-Beltrami and load vector on the QuaQua-mapped surface:
    dict(rowptr, colind, mass, lb, rhs), int64 row pointers"""
    phi = np.ascontiguousarray(phi, np.float64)
    g = np.ascontiguousarray(grad, np.float64).reshape(-1)
    rp = None if rhs_param is None else np.asarray(rhs_param, np.float64)
    P = params if params is not None else quaqua_params()
    L = lib(kind)
    ncut = classify(mesh, phi, zero_policy).size
    cap = 100 * max(ncut, 1)
    rowptr = np.zeros(ndof + 1, np.int64)
    col = np.zeros(cap, np.int32)
    vm, vl = np.zeros(cap), np.zeros(cap)
    rhs = np.zeros(max(ndof, 1))
    nnz = C.c_int64()
    _check(L.orc_assemble_p2qq(C.byref(mesh), _ptr(phi, C.c_double), _ptr(g, C.c_double),
                               _ptr(dof_of_node, C.c_int32), ndof, C.byref(P), rhs_kind, _ptr(rp, C.c_double),
                               zero_policy, cap, C.byref(nnz), _ptr(rowptr, C.c_int64), _ptr(col, C.c_int32),
                               _ptr(vm, C.c_double), _ptr(vl, C.c_double), _ptr(rhs, C.c_double)), "assemble_p2qq")
    k = nnz.value
    return dict(rowptr=rowptr, colind=col[:k].copy(), mass=vm[:k].copy(), lb=vl[:k].copy(), rhs=rhs[:ndof].copy())


def surface_area(mesh, phi, zero_policy=ZERO_ALLOW):
    phi = np.ascontiguousarray(phi, np.float64)
    a = C.c_double()
    _check(lib().orc_surface_area(C.byref(mesh), _ptr(phi, C.c_double), zero_policy, C.byref(a)),
           "surface_area")
    return a.value


def spmv(rowptr, col, val, x):
    y = np.zeros(rowptr.size - 1)
    lib().orc_spmv(rowptr.size - 1, _ptr(rowptr, C.c_int64), _ptr(col, C.c_int32), _ptr(val, C.c_double),
                   _ptr(np.ascontiguousarray(x, np.float64), C.c_double), _ptr(y, C.c_double))
    return y


def kahan_dot(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return lib().orc_kahan_dot(x.size, _ptr(x, C.c_double), _ptr(y, C.c_double))


def pcg_jacobi(rowptr, col, val, b, x0=None, max_iter=10000, tol=1e-10, rel=True, kind="parity"):
    n = rowptr.size - 1
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    it = C.c_int32(max_iter)
    t = C.c_double(tol)
    code = lib(kind).orc_pcg_jacobi(n, _ptr(rowptr, C.c_int64), _ptr(col, C.c_int32), _ptr(val, C.c_double),
                                    _ptr(np.ascontiguousarray(b, np.float64), C.c_double), _ptr(x, C.c_double),
                                    C.byref(it), C.byref(t), int(rel))
    return x, it.value, t.value, code == OK


def gmres_jacobi(rowptr, col, val, b, x0=None, restart=50, max_iter=10000, tol=1e-10, rel=True,
                 left_pc=False, use_pc=True):
    n = rowptr.size - 1
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    it = C.c_int32(max_iter)
    t = C.c_double(tol)
    code = lib().orc_gmres_jacobi(n, _ptr(rowptr, C.c_int64), _ptr(col, C.c_int32), _ptr(val, C.c_double),
                                  _ptr(np.ascontiguousarray(b, np.float64), C.c_double), _ptr(x, C.c_double),
                                  restart, C.byref(it), C.byref(t), int(rel), int(left_pc), int(use_pc))
    return x, it.value, t.value, code == OK


def gmres_diag(rowptr, col, val, pcdiag, b, x0=None, restart=50, max_iter=10000, tol=1e-10, rel=True,
               left_pc=False):
    """Serial GMRES with z = r / pcdiag (block-Jacobi with the given diagonal)."""
    n = rowptr.size - 1
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    it = C.c_int32(max_iter)
    t = C.c_double(tol)
    d = np.ascontiguousarray(pcdiag, np.float64)
    code = lib().orc_gmres_diag(n, _ptr(rowptr, C.c_int64), _ptr(col, C.c_int32), _ptr(val, C.c_double),
                                _ptr(d, C.c_double), _ptr(np.ascontiguousarray(b, np.float64), C.c_double),
                                _ptr(x, C.c_double), restart, C.byref(it), C.byref(t), int(rel), int(left_pc))
    return x, it.value, t.value, code == OK


def tet_mobility(mesh, phi, tet, cv, zero_policy=ZERO_THROW):
    out = np.zeros(16)
    cv = np.ascontiguousarray(cv, np.float64)
    _check(lib().orc_tet_mobility(C.byref(mesh), _ptr(phi, C.c_double), int(tet), _ptr(cv, C.c_double),
                                  zero_policy, _ptr(out, C.c_double)), "tet_mobility")
    return out.reshape(4, 4)


def tet_mobility_p2(mesh, phi, tet, cv, zero_policy=ZERO_THROW):
    out = np.zeros(100)
    cv = np.ascontiguousarray(cv, np.float64)
    _check(lib().orc_tet_mobility_p2(C.byref(mesh), _ptr(phi, C.c_double), int(tet), _ptr(cv, C.c_double),
                                     zero_policy, _ptr(out, C.c_double)), "tet_mobility_p2")
    return out.reshape(10, 10)


def _sfx(space):
    return "_p2" if space == 2 else ""


def ch_assemble(mesh, phi, dof_of_vertex, ndof, rowptr, col, c, zero_policy=ZERO_THROW, kind="parity", space=1):
    """Mobility matrix values on the pattern, M c and M f'(c) (LocalMatVecP1CL accumulations).
    space=2: the P2 interface space, dof_of_vertex is then the half-grid node numbering."""
    lm = np.zeros(col.size)
    td, wl = np.zeros(ndof), np.zeros(ndof)
    c = np.ascontiguousarray(c, np.float64)
    _check(getattr(lib(kind), "orc_ch_assemble" + _sfx(space))(C.byref(mesh), _ptr(phi, C.c_double), _ptr(dof_of_vertex, C.c_int32),
                                     ndof, _ptr(np.ascontiguousarray(rowptr, np.int64), C.c_int64),
                                     _ptr(np.ascontiguousarray(col, np.int32), C.c_int32), _ptr(c, C.c_double),
                                     zero_policy, _ptr(lm, C.c_double), _ptr(td, C.c_double),
                                     _ptr(wl, C.c_double)), "ch_assemble")
    return lm, td, wl


def block2x2(rowptr, col, A, B, Cb, D):
    """SparseMatBaseCL(A, B, C, D) block constructor (num/spmat.h:719-751) for four blocks that
    share one n x n pattern: row i = [A row | B row + n], row n+i = [C row | D row + n]."""
    n = rowptr.size - 1
    rp = np.asarray(rowptr, np.int64)
    ln = np.diff(rp)
    brp = np.zeros(2 * n + 1, np.int64)
    brp[1:] = np.cumsum(np.concatenate([2 * ln, 2 * ln]))
    bcol = np.empty(4 * col.size, np.int32)
    bval = np.empty(4 * col.size)
    for i in range(n):
        s, e = rp[i], rp[i + 1]
        for I, (L, R) in ((i, (A, B)), (n + i, (Cb, D))):
            o = brp[I]
            k = e - s
            bcol[o:o + k] = col[s:e]
            bcol[o + k:o + 2 * k] = col[s:e] + n
            bval[o:o + k] = L[s:e]
            bval[o + k:o + 2 * k] = R[s:e]
    return brp, bcol, bval


def ch_system(mesh, phi, dof_of_vertex, ndof, rowptr, col, mass, lb, stab, c, dt, eps, rho, sigma=1.0, S=1.0,
              kind="parity", space=1):
    """Block system of one step (orc_ch_system): block CSR, rhs and block-Jacobi diagonal."""
    nnz = col.size
    brp = np.zeros(2 * ndof + 1, np.int64)
    bcol = np.zeros(4 * nnz, np.int32)
    bval, rhs, pcd = np.zeros(4 * nnz), np.zeros(2 * ndof), np.zeros(2 * ndof)
    f64 = lambda a: _ptr(np.ascontiguousarray(a, np.float64), C.c_double)
    _check(getattr(lib(kind), "orc_ch_system" + _sfx(space))(C.byref(mesh), _ptr(phi, C.c_double), _ptr(dof_of_vertex, C.c_int32), ndof,
                                   _ptr(np.ascontiguousarray(rowptr, np.int64), C.c_int64),
                                   _ptr(np.ascontiguousarray(col, np.int32), C.c_int32), f64(mass), f64(lb), f64(stab),
                                   f64(c), dt, eps, rho, sigma, S, _ptr(brp, C.c_int64), _ptr(bcol, C.c_int32),
                                   _ptr(bval, C.c_double), _ptr(rhs, C.c_double), _ptr(pcd, C.c_double)), "ch_system")
    return brp, bcol, bval, rhs, pcd


def ch_step_c(mesh, phi, dof_of_vertex, ndof, rowptr, col, mass, lb, stab, x, dt, eps, rho, sigma=1.0, S=1.0,
              restart=50, tol=1e-8, max_iter=20000, rel=True, kind="parity", space=1):
    """One full step in C (orc_ch_step); x = [mu; c] updated in place.  Returns (iters, resid, ok)."""
    it, t = C.c_int32(max_iter), C.c_double(tol)
    f64 = lambda a: _ptr(np.ascontiguousarray(a, np.float64), C.c_double)
    code = getattr(lib(kind), "orc_ch_step" + _sfx(space))(C.byref(mesh), _ptr(phi, C.c_double), _ptr(dof_of_vertex, C.c_int32), ndof,
                                 _ptr(np.ascontiguousarray(rowptr, np.int64), C.c_int64),
                                 _ptr(np.ascontiguousarray(col, np.int32), C.c_int32), f64(mass), f64(lb), f64(stab),
                                 _ptr(x, C.c_double), dt, eps, rho, sigma, S, restart, C.byref(it), C.byref(t),
                                 int(rel))
    if code not in (OK, E_NOT_CONVERGED):
        _check(code, "ch_step")
    return it.value, t.value, code == OK


def ch_step(mesh, phi, dof_of_vertex, ndof, rowptr, col, mass, lb, stab, c, mu, dt, eps, rho,
            sigma=1.0, S=1.0, restart=50, tol=1e-8, max_iter=20000, kind="parity", space=1):
    """One step of CahnHilliardcGP1CL::DoStep (surfactant/ifacetransp.cpp:1823-1856) with theta = 1,
    resting interface (no wind: Conv = Massd = 0) and no sources, solved by right-preconditioned
    GMRES(restart) with block Jacobi: the diagonals of Precond3 = M + sigma dt LM + dt rho C and
    Precond4 = M + eps^2 L + eps^2 rho C (:1567-1570).  Unknowns [mu; c], start = previous step.
    The system comes from orc_ch_system; the mobility matrix and LocalMatVec vectors are returned
    too (orc_ch_assemble)."""
    lm, timeder, well = ch_assemble(mesh, phi, dof_of_vertex, ndof, rowptr, col, c, kind=kind, space=space)
    brp, bcol, bval, b, pcd = ch_system(mesh, phi, dof_of_vertex, ndof, rowptr, col, mass, lb, stab, c, dt, eps,
                                        rho, sigma, S, kind=kind, space=space)
    x0 = np.concatenate([mu, c])
    x, it, res, ok = gmres_diag(brp, bcol, bval, pcd, b, x0, restart=restart, max_iter=max_iter, tol=tol)
    return dict(mu=x[:ndof], c=x[ndof:], iters=it, res=res, ok=ok, lm=lm, timeder=timeder, well=well,
                rhs=b, pcdiag=pcd, rowptr=brp, col=bcol, val=bval, x0=x0)


def bdf2_weights(dt, dt_prev, kind="parity"):
    """(dt_eff, w1, w2, w3, w4) of orc_bdf2_weights: c_e = w1 c_n + w2 c_{n-1}, c_r = w3 c_n + w4 c_{n-1}."""
    w = np.zeros(5)
    lib(kind).orc_bdf2_weights(dt, dt_prev, _ptr(w, C.c_double))
    return tuple(w)


def bdf_params(step_size, tol, rho=0.9, min_step=0.0, max_step=float("inf"), adaptive=True, max_tries=0):
    return BdfParams(step_size, tol, rho, min_step, max_step, int(adaptive), max_tries)


def bdf_step_size(params, e, dt, kind="parity"):
    """F(e, dt) of surfphasesep/surfphasesep.cpp:117-123."""
    return lib(kind).orc_bdf_step_size(C.byref(params), e, dt)


def ch_system_bdf(mesh, phi, dof_of, ndof, rowptr, col, mass, lb, stab, c_e, c_r, dt_eff, eps, rho, sigma=1.0,
                  S=1.0, kind="parity", space=1):
    """BDF2 step system (orc_ch_system_bdf): mobility and f' from c_e, rhs3 = M c_r, effective step dt_eff."""
    nnz = col.size
    brp = np.zeros(2 * ndof + 1, np.int64)
    bcol = np.zeros(4 * nnz, np.int32)
    bval, rhs, pcd = np.zeros(4 * nnz), np.zeros(2 * ndof), np.zeros(2 * ndof)
    f64 = lambda a: _ptr(np.ascontiguousarray(a, np.float64), C.c_double)
    _check(lib(kind).orc_ch_system_bdf(space, C.byref(mesh), _ptr(phi, C.c_double), _ptr(dof_of, C.c_int32), ndof,
                                       _ptr(np.ascontiguousarray(rowptr, np.int64), C.c_int64),
                                       _ptr(np.ascontiguousarray(col, np.int32), C.c_int32), f64(mass), f64(lb),
                                       f64(stab), f64(c_e), f64(c_r), dt_eff, eps, rho, sigma, S,
                                       _ptr(brp, C.c_int64), _ptr(bcol, C.c_int32), _ptr(bval, C.c_double),
                                       _ptr(rhs, C.c_double), _ptr(pcd, C.c_double)), "ch_system_bdf")
    return brp, bcol, bval, rhs, pcd


class BdfStepper:
    """The adaptive BDF1/BDF2 time loop of surfphasesep/surfphasesep.cpp:396-491 on the step system
    of the Cahn-Hilliard stepper (orc_ch_bdf_step).  x = [mu; c] is the accepted state."""

    def __init__(self, mesh, phi, dof_of, ndof, r, c0, mu0, params, eps, rho, sigma=1.0, S=1.0, restart=50,
                 max_iter=20000, tol=1e-10, rel=True, kind="parity", space=1):
        self.args = (mesh, phi, dof_of, ndof, np.ascontiguousarray(r["rowptr"], np.int64),
                     np.ascontiguousarray(r["colind"], np.int32), np.ascontiguousarray(r["mass"], np.float64),
                     np.ascontiguousarray(r["lb"], np.float64), np.ascontiguousarray(r["stab"], np.float64))
        self.n = ndof
        self.x = np.concatenate([np.asarray(mu0, np.float64), np.asarray(c0, np.float64)])
        self.cprev = np.zeros(ndof)
        self.state = BdfState(0, 0.0, 0.0, 0.0)
        self.params = params
        self.phys = (eps, rho, sigma, S)
        self.solver = (restart, max_iter, tol, int(rel))
        self.kind, self.space = kind, space

    @property
    def c(self):
        return self.x[self.n:]

    @property
    def mu(self):
        return self.x[:self.n]

    def step(self):
        """One time step; returns dict(step, t, dt, e, tries, iters_bdf1, iters_bdf2, resid)."""
        mesh, phi, dof, nd, rp, col, mass, lb, stab = self.args
        st = BdfStat()
        _check(lib(self.kind).orc_ch_bdf_step(self.space, C.byref(mesh), _ptr(phi, C.c_double),
                                              _ptr(dof, C.c_int32), nd, _ptr(rp, C.c_int64),
                                              _ptr(col, C.c_int32), _ptr(mass, C.c_double),
                                              _ptr(lb, C.c_double), _ptr(stab, C.c_double),
                                              _ptr(self.x, C.c_double), _ptr(self.cprev, C.c_double),
                                              C.byref(self.state), C.byref(self.params), *self.phys,
                                              *self.solver, C.byref(st)), "ch_bdf_step")
        return dict(step=self.state.step, t=self.state.t, dt=self.state.dt, e=self.state.e, tries=st.tries,
                    iters_bdf1=st.iters_bdf1, iters_bdf2=st.iters_bdf2, resid=st.resid)


# ---------------------------------------------------------------- file formats (iostreams)
_fmt_lib = None


def fmt_lib():
    global _fmt_lib
    if _fmt_lib is None:
        L = C.CDLL(os.path.join(HERE, "_build", "libformat_oracle.so"))
        I64p, Dp, CP = C.POINTER(C.c_int64), C.POINTER(C.c_double), C.c_char_p
        L.orc_fmt_mtx.argtypes = [CP, C.c_int64, C.c_int64, I64p, I64p, Dp]
        L.orc_fmt_plain.argtypes = [CP, C.c_int64, C.c_int64, I64p, I64p, Dp]
        L.orc_read_plain.argtypes = [CP, I64p, I64p, I64p, I64p, I64p, Dp]
        L.orc_fmt_vec.argtypes = [CP, Dp, C.c_int64, C.c_int]
        L.orc_read_vec.argtypes = [CP, Dp, C.c_int64, C.c_int]
        _fmt_lib = L
    return _fmt_lib


def fmt_matrix(path, rowptr, colind, val, ncols=None, fmt="mtx"):
    """operator<< (fmt "mtx", the body of exportMTX) or out() (fmt "plain") through iostreams"""
    rp = np.ascontiguousarray(rowptr, np.int64)
    ci = np.ascontiguousarray(colind, np.int64)
    v = np.ascontiguousarray(val, np.float64)
    n = rp.size - 1
    f = fmt_lib().orc_fmt_mtx if fmt == "mtx" else fmt_lib().orc_fmt_plain
    _check(f(os.fsencode(path), n, n if ncols is None else ncols, _ptr(rp, C.c_int64), _ptr(ci, C.c_int64),
             _ptr(v, C.c_double)), "fmt_matrix")


def read_plain(path):
    """in(): dict(shape, rowptr, colind, val)"""
    nr, nc, nz = C.c_int64(), C.c_int64(), C.c_int64()
    b = os.fsencode(path)
    _check(fmt_lib().orc_read_plain(b, C.byref(nr), C.byref(nc), C.byref(nz), None, None, None), "read_plain")
    rp = np.zeros(nr.value + 1, np.int64)
    ci = np.zeros(max(nz.value, 1), np.int64)
    v = np.zeros(max(nz.value, 1))
    _check(fmt_lib().orc_read_plain(b, C.byref(nr), C.byref(nc), C.byref(nz), _ptr(rp, C.c_int64),
                                    _ptr(ci, C.c_int64), _ptr(v, C.c_double)), "read_plain")
    return dict(shape=(nr.value, nc.value), rowptr=rp, colind=ci[:nz.value], val=v[:nz.value])


def fmt_vector(path, x, binary=True):
    a = np.ascontiguousarray(x, np.float64)
    _check(fmt_lib().orc_fmt_vec(os.fsencode(path), _ptr(a, C.c_double), a.size, int(binary)), "fmt_vector")


def read_vector(path, n, binary=True):
    """VecDescBaseCL::Read; raises OracleError(1) when the stored count differs from n"""
    out = np.zeros(max(n, 1))
    _check(fmt_lib().orc_read_vec(os.fsencode(path), _ptr(out, C.c_double), n, int(binary)), "read_vector")
    return out[:n]


# ---------------------------------------------------------------- narrow band
def classify_band(mesh, phi, width, zero_policy=ZERO_THROW):
    """(band tets in mesh order, is_cut flags) of distance() <= width"""
    phi = np.ascontiguousarray(phi, np.float64)
    n = C.c_int64()
    _check(lib().orc_classify_band(C.byref(mesh), _ptr(phi, C.c_double), width, zero_policy, 0, C.byref(n), None,
                                   None), "classify_band")
    t = np.zeros(max(n.value, 1), np.int64)
    f = np.zeros(max(n.value, 1), np.uint8)
    _check(lib().orc_classify_band(C.byref(mesh), _ptr(phi, C.c_double), width, zero_policy, n.value, C.byref(n),
                                   _ptr(t, C.c_int64), _ptr(f, C.c_uint8)), "classify_band")
    return t[:n.value], f[:n.value]


def number_p1band(mesh, phi, width, zero_policy=ZERO_THROW):
    phi = np.ascontiguousarray(phi, np.float64)
    nv = (mesh.n[0] + 1) * (mesh.n[1] + 1) * (mesh.n[2] + 1)
    dof = np.empty(nv, np.int32)
    nd = C.c_int64()
    _check(lib().orc_number_p1band(C.byref(mesh), _ptr(phi, C.c_double), width, zero_policy, _ptr(dof, C.c_int32),
                                   C.byref(nd)), "number_p1band")
    return dof, nd.value


def assemble_p1band(mesh, phi, dof, ndof, width, coef=(1.0, 1.0, 1.0), rhs_kind=RHS_NONE, rhs_param=None,
                    zero_policy=ZERO_THROW):
    """dict(stab=(rowptr, colind, val), surf=(rowptr, colind, mass, lb), rhs): the stabilisation on its
    own (band) pattern, mass and LB on theirs (cut tets)"""
    phi = np.ascontiguousarray(phi, np.float64)
    dof = np.ascontiguousarray(dof, np.int32)
    co = np.asarray(coef, np.float64)
    rp = None if rhs_param is None else np.asarray(rhs_param, np.float64)
    cap = 27 * max(ndof, 1)
    rps, rpm = np.zeros(ndof + 1, np.int64), np.zeros(ndof + 1, np.int64)
    cs, cm = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    vs, vm, vl = np.zeros(cap), np.zeros(cap), np.zeros(cap)
    rhs = np.zeros(max(ndof, 1))
    ns, nm = C.c_int64(), C.c_int64()
    _check(lib().orc_assemble_p1band(C.byref(mesh), _ptr(phi, C.c_double), _ptr(dof, C.c_int32), ndof, width,
                                     _ptr(co, C.c_double), rhs_kind, _ptr(rp, C.c_double), zero_policy, cap,
                                     C.byref(ns), _ptr(rps, C.c_int64), _ptr(cs, C.c_int32), _ptr(vs, C.c_double),
                                     C.byref(nm), _ptr(rpm, C.c_int64), _ptr(cm, C.c_int32), _ptr(vm, C.c_double),
                                     _ptr(vl, C.c_double), _ptr(rhs, C.c_double)), "assemble_p1band")
    return dict(stab=(rps, cs[:ns.value], vs[:ns.value]), surf=(rpm, cm[:nm.value], vm[:nm.value], vl[:nm.value]),
                rhs=rhs[:ndof])
