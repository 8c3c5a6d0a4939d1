"""GPU parity of the P2 forms on the QuaQua-mapped surface (dhip_assemble_p2_quaqua, dhip_local_p2_quaqua)
against the oracle, through the C ABI.

Bar: the local 10x10 mass and Laplace-Beltrami forms and the load vector, the CSR pattern (only tets with a
lattice-1 patch add couplings) and every assembled entry bit-exact against the oracle (both sides run the same
Newton iterations, QR factorisations and sums in the same order without FMA contraction) where the tet's
cube root of |det| is exact (mesh width a power of two); elsewhere within 1e-12 relative, since cbrt may round
differently there and it only enters the Newton damping length.  At 128^3 (the surface-solve size): the mapped
sphere's mass sums to 4 pi (to what the Newton tolerance allows), the Laplace-Beltrami rows to zero, and
the load of f = 1 is M 1.  A load given as values at the base points the library lists
(dhip_quaqua_points / dhip_quaqua_values) equals the device kind's bit for bit, is linear in f, and is
refused when stale.  Calls out of order, bad parameters and exact zeros of the level set at nodes of cut
tets are refused.  Ragged, sheared and column bricks, a plane (whose mapped surface is itself) and a single
cell match the oracle; an interface-free level set maps and assembles nothing."""
import numpy as np
import pytest

import cases
import oracle as O
from drops_amd import dhip as D

pytestmark = pytest.mark.gpu

LC = (1.0, 0.5, 0.0)


def run(ctx, mesh, phi, rhs=D.RHS_3XY_MINUS_X3, rhs_param=None, params=None):
    print(f"[gpu quaqua] mesh n={mesh.n}", flush=True)
    ctx.mesh(mesh)
    ctx.levelset(phi)
    return ctx.assemble_p2_quaqua(params=params, rhs_kind=rhs, rhs_param=rhs_param, lc=LC)


@pytest.mark.parametrize("case,n,exact", [(cases.c1_sphere, 16, True), (cases.torus_case, 24, False),
                                          (cases.c1_sphere, 32, True)])
def test_mapped_forms_match_oracle(ctx, case, n, exact):
    mesh, phi = case(n)
    ncut, ndof, nnz = run(ctx, mesh, phi)
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    grad = O.recovered_gradient(om, phi)
    cut = ctx.cut_tets()
    loc, flags = ctx.local_p2()
    worst = 0.0
    for c in range(0, ncut, max(1, ncut // 400)):
        r = O.tet_local_p2qq(om, phi, grad, int(cut[c]), rhs_kind=O.RHS_3XY_MINUS_X3)
        assert bool(flags[c] & 1) == (r.nfacet > 0)
        ref = np.concatenate([np.ctypeslib.as_array(r.mass), np.ctypeslib.as_array(r.lb),
                              np.zeros(100), np.ctypeslib.as_array(r.rhs)])
        got = loc[c, :310]
        if exact:
            assert np.array_equal(got, ref), f"tet {cut[c]}"
        else:
            worst = max(worst, float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)))
    assert worst <= 1e-12
    dof, nd = O.number_p2if(om, phi)
    assert ndof == nd
    ref = O.assemble_p2qq(om, phi, grad, dof, nd, rhs_kind=O.RHS_3XY_MINUS_X3)
    g = ctx.csr()
    assert nnz == ref["colind"].size
    assert np.array_equal(g["rowptr"].astype(np.int64), ref["rowptr"])
    assert np.array_equal(g["colind"], ref["colind"])
    for k in ("mass", "lb", "rhs"):
        if exact:
            assert np.array_equal(g[k], ref[k]), k
        else:
            assert np.abs(g[k] - ref[k]).max() <= 1e-12 * np.abs(ref[k]).max(), k
    assert not g["stab"].any()
    if exact:
        assert np.array_equal(g["comb"], 1.0 * (LC[0] * ref["mass"] + LC[1] * ref["lb"]) + LC[2] * 0.0)


@pytest.mark.parametrize("tol,bar", [(1e-7, 5e-6), (1e-12, 1e-10)])
def test_mapped_sphere_at_surface_solve_size(ctx, tol, bar):
    # The Newton stops at |F| < tol, |F| >= |phi_h(y)|: with the reference's 1e-7 the base points sit up to
    # ~1e-7 off the sphere, which bounds the area's accuracy on fine meshes (96^3: 1.4e-7, against 2.4e-11
    # at tol 1e-12 on the oracle).
    mesh, phi = cases.c1_sphere(128)
    ncut, ndof, nnz = run(ctx, mesh, phi, rhs=D.RHS_CONST, rhs_param=[1.0], params=D.QuaQuaParams(tol=tol))
    g = ctx.csr()
    rp = g["rowptr"].astype(np.int64)
    rows = np.repeat(np.arange(ndof), np.diff(rp))
    area = g["mass"].sum()
    assert abs(area - 4 * np.pi) < bar, area - 4 * np.pi
    scale = np.abs(g["lb"]).max()
    lbsum = np.bincount(rows, weights=g["lb"], minlength=ndof)
    assert np.abs(lbsum).max() <= 1e-12 * scale
    msum = np.bincount(rows, weights=g["mass"], minlength=ndof)
    assert np.abs(g["rhs"] - msum).max() <= 1e-15
    # symmetric pattern and values
    import scipy.sparse as sp
    for k in ("mass", "lb"):
        A = sp.csr_matrix((g[k], g["colind"], rp), shape=(ndof, ndof))
        assert abs(A - A.T).max() <= 1e-15 * abs(A).max(), k


@pytest.mark.parametrize("name", ["ellipsoid_ragged", "skew_brick", "plane_x041", "sphere20"])
def test_mapped_forms_other_geometries(ctx, name):
    """Ragged brick (17x13x11), sheared parallelepiped brick, a plane through a 10x1x1 column, the unit-cube
    sphere of tests/interfaceP1FE.cpp: the CSR pattern bit-exact against the oracle, every value within
    1e-12 relative; the plane's mapped surface is the plane itself (area 1)."""
    if name == "plane_x041":
        mesh, phi, _ = cases.planes()
    else:
        mesh, phi = cases.PARITY_CASES[name]()
    ncut, ndof, nnz = run(ctx, mesh, phi, rhs=D.RHS_CONST, rhs_param=[1.0])
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    grad = O.recovered_gradient(om, phi)
    dof, nd = O.number_p2if(om, phi)
    ref = O.assemble_p2qq(om, phi, grad, dof, nd, rhs_kind=O.RHS_CONST, rhs_param=[1.0])
    g = ctx.csr()
    assert ndof == nd and nnz == ref["colind"].size
    assert np.array_equal(g["rowptr"].astype(np.int64), ref["rowptr"])
    assert np.array_equal(g["colind"], ref["colind"])
    for k in ("mass", "lb", "rhs"):
        assert np.abs(g[k] - ref[k]).max() <= 1e-12 * np.abs(ref[k]).max(), k
    if name == "plane_x041":
        assert abs(g["mass"].sum() - 1.0) <= 1e-14
        gr = ctx.recovered_gradient()   # the gradient of x - 0.41, up to rounding
        assert np.abs(gr - np.array([1.0, 0.0, 0.0])).max() <= 1e-12


def test_mapped_single_cell_and_empty(ctx):
    """One brick cell with one negative corner (the oracle's numbering and matrices), and an interface-free
    level set: nothing to map, nothing assembled, no base points."""
    from drops_amd.mesh import KuhnMesh
    mesh = KuhnMesh.box(0.0, 1.0, 1)
    phi = np.ones(mesh.halfgrid_shape)
    phi[0, 0, 2] = -1.0
    ncut, ndof, nnz = run(ctx, mesh, phi, rhs=D.RHS_CONST, rhs_param=[1.0])
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    grad = O.recovered_gradient(om, phi)
    dof, nd = O.number_p2if(om, phi)
    ref = O.assemble_p2qq(om, phi, grad, dof, nd, rhs_kind=O.RHS_CONST, rhs_param=[1.0])
    g = ctx.csr()
    assert ncut == 2 and ndof == nd and nnz == ref["colind"].size
    assert np.array_equal(g["colind"], ref["colind"])
    for k in ("mass", "lb", "rhs"):
        assert np.abs(g[k] - ref[k]).max() <= 1e-12 * max(np.abs(ref[k]).max(), 1e-300), k
    phi[:] = 1.0
    assert run(ctx, mesh, phi) == (0, 0, 0)
    assert ctx.csr()["rowptr"].tolist() == [0]
    ctx.classify()
    assert ctx.quaqua_points().shape == (0, 3)
    ctx.quaqua_values(np.zeros(0))
    assert ctx.assemble_p2_quaqua(rhs_kind=D.RHS_VALUES, lc=LC) == (0, 0, 0)


def test_mapped_calls_refused(ctx):
    mesh, phi = cases.c1_sphere(12)
    ctx.mesh(mesh)
    ctx.levelset(phi)
    ctx.classify()
    with pytest.raises(D.DhipError) as ei:
        ctx.check(ctx.L.dhip_local_p2_quaqua(ctx.h, None, D.RHS_NONE, None))
    assert ei.value.code == D.E_STATE
    with pytest.raises(D.DhipError) as ei:   # values before any base points
        ctx.assemble_p2_quaqua(rhs_kind=D.RHS_VALUES)
    assert ei.value.code == D.E_STATE
    with pytest.raises(D.DhipError) as ei:
        ctx.assemble_p2_quaqua(rhs_kind=99)
    assert ei.value.code == D.E_ARG
    with pytest.raises(D.DhipError) as ei:
        ctx.assemble_p2_quaqua(params=D.QuaQuaParams(maxiter=-1))
    assert ei.value.code == D.E_ARG
    # a fresh, valid call after the refusals
    ncut, ndof, nnz = ctx.assemble_p2_quaqua()
    assert ndof > 0 and nnz > 0
    # a level set exactly zero at nodes of cut tets (the plane x = 0.4 through a layer of vertices): the
    # reference's numbering throws, and so does the mapped pass under DHIP_ZERO_THROW
    pm, _, p04 = cases.planes()
    ctx.mesh(pm)
    ctx.levelset(p04)
    with pytest.raises(D.DhipError) as ei:
        ctx.assemble_p2_quaqua()
    assert ei.value.code == D.E_ZERO_NODE


def f_kind(xyz):
    """f = 3xy - x^3 with the device's evaluation order: (3 x) y - (x x) x"""
    x, y = xyz[:, 0], xyz[:, 1]
    return (3.0 * x) * y - (x * x) * x


@pytest.mark.parametrize("case,n", [(cases.c1_sphere, 24), (cases.torus_case, 16)])
def test_mapped_load_from_values(ctx, case, n):
    mesh, phi = case(n)
    print(f"[gpu quaqua] values n={n}", flush=True)
    P = D.QuaQuaParams()
    run(ctx, mesh, phi, params=P)
    ref = ctx.csr()["rhs"].copy()
    run(ctx, mesh, phi, rhs=D.RHS_CONST, rhs_param=[2.5], params=P)
    cref = ctx.csr()["rhs"].copy()
    flags = ctx.local_p2()[1]
    ctx.classify()
    xyz = ctx.quaqua_points(P)
    # 7 points per facet of each cut tet's lattice-1 patch, in the cut tets' order
    assert xyz.shape[0] > 0 and xyz.shape[0] % 7 == 0
    assert (flags & 1).sum() <= xyz.shape[0] // 7 <= 2 * (flags & 1).sum()
    if case is cases.c1_sphere:   # the base points lie on the sphere, to the Newton tolerance
        assert np.abs(np.linalg.norm(xyz, axis=1) - 1.0).max() <= 1e-7
    # the points do not depend on the call: a second listing is identical
    assert np.array_equal(ctx.quaqua_points(P), xyz)
    ctx.quaqua_values(f_kind(xyz))
    ctx.assemble_p2_quaqua(params=P, rhs_kind=D.RHS_VALUES, lc=LC)
    assert np.array_equal(ctx.csr()["rhs"], ref)
    ctx.quaqua_values(np.full(xyz.shape[0], 2.5))
    ctx.assemble_p2_quaqua(params=P, rhs_kind=D.RHS_VALUES, lc=LC)
    assert np.array_equal(ctx.csr()["rhs"], cref)
    g1, g2 = np.sin(xyz[:, 0]) * xyz[:, 2], np.cos(3.0 * xyz[:, 1])
    b = []
    for g in (g1, g2, g1 + g2):
        ctx.quaqua_values(g)
        ctx.assemble_p2_quaqua(params=P, rhs_kind=D.RHS_VALUES, lc=LC)
        b.append(ctx.csr()["rhs"].copy())
    assert np.abs(b[2] - (b[0] + b[1])).max() <= 1e-14 * (np.abs(b[0]).max() + np.abs(b[1]).max())


def test_mapped_values_refused_when_stale(ctx):
    mesh, phi = cases.c1_sphere(16)
    print("[gpu quaqua] values state checks", flush=True)
    ctx.mesh(mesh)
    ctx.levelset(phi)
    ctx.classify()
    xyz = ctx.quaqua_points()
    with pytest.raises(D.DhipError) as ei:
        ctx.quaqua_values(np.zeros(xyz.shape[0] + 1))
    assert ei.value.code == D.E_ARG
    ctx.quaqua_values(f_kind(xyz))
    with pytest.raises(D.DhipError) as ei:   # other Newton parameters: other base points
        ctx.assemble_p2_quaqua(params=D.QuaQuaParams(tol=1e-12), rhs_kind=D.RHS_VALUES)
    assert ei.value.code == D.E_STATE
    ctx.levelset(phi)          # a new level set: the values no longer belong to the cut tets
    with pytest.raises(D.DhipError) as ei:
        ctx.assemble_p2_quaqua(rhs_kind=D.RHS_VALUES)
    assert ei.value.code == D.E_STATE
    with pytest.raises(D.DhipError) as ei:
        ctx.quaqua_values(f_kind(xyz))
    assert ei.value.code == D.E_STATE


@pytest.mark.parametrize("case,n", [(cases.c1_sphere, 12), (cases.torus_case, 16)])
def test_recovered_gradient_matches_oracle(ctx, case, n):
    mesh, phi = case(n)
    print(f"[gpu quaqua] recovered gradient n={n}", flush=True)
    ctx.mesh(mesh)
    ctx.levelset(phi)
    g = ctx.recovered_gradient()
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    assert np.array_equal(g, O.recovered_gradient(om, phi))


def test_host_mirror_mapped_forms(tmp_path):
    """The reference's set-up (surfactant/surfactant.cpp:1596-1640: recovered gradient, neighbourhoods,
    QuaQuaMapperCL, InterfaceCommonDataP2CL on lattice 1 and the three accumulators) through the C++ mirror
    gives the oracle's matrices and load bit for bit."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "drops_amd", "lib", "tracefem_p1")
    print("[gpu quaqua] host mirror", flush=True)
    r = subprocess.run([exe, "quaqua", "16", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rd = lambda name, dt: np.fromfile(tmp_path / name, dt)
    mesh, phi = cases.c1_sphere(16)
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    grad = O.recovered_gradient(om, phi)
    dof, nd = O.number_p2if(om, phi)
    ref = O.assemble_p2qq(om, phi, grad, dof, nd, rhs_kind=O.RHS_3XY_MINUS_X3)
    assert f"P2-NumUnknowns: {nd}" in r.stdout
    assert np.array_equal(rd("rowptr.bin", np.int64), ref["rowptr"])
    assert np.array_equal(rd("colind.bin", np.int64), ref["colind"].astype(np.int64))
    for k in ("mass", "lb", "rhs"):
        assert np.array_equal(rd(f"{k}.bin", np.float64), ref[k]), k
    assert np.array_equal(rd("grad.bin", np.float64), grad.reshape(-1))
    assert abs(float(r.stdout.split("mapped area - 4 pi:")[1].split()[0])) < 3e-6
    # LocalVectorP2CL with a host function: f at the listed base points, the device kind's load bit for bit
    assert float(r.stdout.split("user rhs vs device rhs:")[1].split()[0]) == 0.0
