"""GPU parity of the trace-FEM assembly against the CPU oracle, through the C ABI.

Bar: cut classification, DOF numbering and the CSR pattern bit-exact; local
matrices, assembled entries and load vector bit-exact as well (both sides
evaluate every product and sum in the reference's order without FMA
contraction), which implies the 1e-12 relative bar of the path.
"""
import numpy as np
import pytest

import cases
import oracle as O
from drops_amd import dhip as D

pytestmark = pytest.mark.gpu

KUHN = np.array([[[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, 1]],
                 [[0, 0, 0], [1, 0, 0], [1, 0, 1], [1, 1, 1]],
                 [[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 1, 1]],
                 [[0, 0, 0], [0, 1, 0], [0, 1, 1], [1, 1, 1]],
                 [[0, 0, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1]],
                 [[0, 0, 0], [0, 0, 1], [0, 1, 1], [1, 1, 1]]])


def tet_vertex_ids(mesh, tets):
    n1, n2, _ = mesh.n
    cell, k = tets // 6, tets % 6
    i1, i2, i3 = cell % n1, (cell // n1) % n2, cell // (n1 * n2)
    off = KUHN[k]  # (T,4,3)
    v1, v2, v3 = i1[:, None] + off[..., 0], i2[:, None] + off[..., 1], i3[:, None] + off[..., 2]
    return (v3 * (n2 + 1) + v2) * (n1 + 1) + v1


def run_gpu(ctx, mesh, phi, coef=(1.0, 0.7, 1.3), rhs=D.RHS_3XY_MINUS_X3, lc=(1.0, 1.0, 0.1), zero=D.ZERO_THROW):
    print(f"[gpu assembly] mesh n={mesh.n} coef={coef}", flush=True)
    ctx.mesh(mesh)
    ctx.levelset(phi)
    ncut, ndof, nnz = ctx.assemble_p1(zero_policy=zero, coef=coef, rhs_kind=rhs, lc=lc)
    return ncut, ndof, nnz


@pytest.mark.parametrize("name", sorted(cases.PARITY_CASES))
def test_assembly_bit_exact(ctx, name):
    mesh, phi = cases.PARITY_CASES[name]()
    coef, lc = (1.0, 0.7, 1.3), (1.0, 1.0, 0.1)
    ncut, ndof, nnz = run_gpu(ctx, mesh, phi, coef=coef, lc=lc)
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    # classification
    cut_ref = O.classify(om, phi)
    cut = ctx.cut_tets()
    assert np.array_equal(cut, cut_ref)
    # local matrices of every cut tet
    loc, flags = ctx.local()
    assert np.all(flags & 1)
    for c in range(ncut):
        r = O.tet_local(om, phi, int(cut[c]), coef=coef, rhs_kind=O.RHS_3XY_MINUS_X3)
        ref = np.concatenate([np.ctypeslib.as_array(r.mass), np.ctypeslib.as_array(r.lb),
                              np.ctypeslib.as_array(r.stab), np.ctypeslib.as_array(r.rhs),
                              np.ctypeslib.as_array(r.hat_sq), [r.area]])
        if not np.array_equal(loc[c], ref):
            d = np.abs(loc[c] - ref) / np.maximum(np.abs(ref), 1e-300)
            raise AssertionError(f"tet {cut[c]}: local data differ, max rel {d.max():.3e} at {int(d.argmax())}")
    # numbering
    dof_ref, nd_ref = O.number_p1if(om, phi)
    assert ndof == nd_ref
    dofs = ctx.tet_dofs()
    vid = tet_vertex_ids(mesh, cut)
    dof_gpu = np.full(mesh.num_vertices, -1, np.int64)
    dof_gpu[vid.ravel()] = dofs.ravel()
    assert np.array_equal(dof_gpu, dof_ref.astype(np.int64))
    vod = ctx.dof_vertices()
    assert np.array_equal(dof_ref[vod], np.arange(ndof))
    # CSR
    g = ctx.csr()
    ref = O.assemble_p1(om, phi, dof_ref, nd_ref, coef=coef, rhs_kind=O.RHS_3XY_MINUS_X3)
    assert nnz == ref["colind"].size
    assert np.array_equal(g["rowptr"].astype(np.int64), ref["rowptr"])
    assert np.array_equal(g["colind"], ref["colind"])
    for k in ("mass", "lb", "stab", "rhs"):
        assert np.array_equal(g[k], ref[k]), k
    comb = (1.0 * (lc[0] * ref["mass"] + lc[1] * ref["lb"])) + lc[2] * ref["stab"]
    assert np.array_equal(g["comb"], comb)


def test_known_answer_sphere_dofs(ctx):
    """TestinterfaceP1FE.ref: (sphere) NumUnknowns 1156, through the GPU path."""
    mesh, phi = cases.sphere20()
    _, ndof, _ = run_gpu(ctx, mesh, phi)
    assert ndof == 1156


def test_known_answer_surface_area(ctx):
    """Testprincipallattice.ref: surface 3.0909 (rel 1e-4); exact zeros on the grid -> zero-allow mode."""
    mesh, phi = cases.area_case()
    ctx.mesh(mesh)
    ctx.levelset(phi)
    ncut = ctx.classify(D.ZERO_ALLOW)
    ctx.local_p1()
    loc, flags = ctx.local()
    area = 0.0
    for a in loc[:, 56]:
        area += a  # per-tet sums added in mesh order, as the reference test
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    assert area == O.surface_area(om, phi)
    assert abs(area - 3.0909) / 3.0909 < 1e-4
    assert ncut > 0


def test_planes_and_zero_nodes(ctx):
    """TestinterfaceP1FE.ref: planar interface 8 DOFs; the plane through grid vertices
    raises the reference's exact-zero error, and gives the historical 4 DOFs when zeros are allowed."""
    mesh, p041, p040 = cases.planes()
    assert run_gpu(ctx, mesh, p041)[1] == 8
    with pytest.raises(D.DhipError) as ei:
        run_gpu(ctx, mesh, p040)
    assert ei.value.code == D.E_ZERO_NODE
    assert run_gpu(ctx, mesh, p040, zero=D.ZERO_ALLOW)[1] == 4


def test_empty_interface(ctx):
    """No cut tets: empty numbering and matrix."""
    from drops_amd.mesh import KuhnMesh
    mesh = KuhnMesh.box(-1.0, 1.0, (5, 4, 3))
    phi = mesh.sample_p2(lambda x, y, z: x * x + y * y + z * z + 1.0)
    assert run_gpu(ctx, mesh, phi) == (0, 0, 0)
    g = ctx.csr()
    assert g["rowptr"].tolist() == [0]


def test_single_cell(ctx):
    """One brick cell with one negative corner (cf. the single-tetra case of tests/interfaceP1FE.cpp:33-61)."""
    from drops_amd.mesh import KuhnMesh
    mesh = KuhnMesh.box(0.0, 1.0, 1)
    phi = np.ones(mesh.halfgrid_shape)
    phi[0, 0, 2] = -1.0  # vertex (1,0,0): in Kuhn tets 0 and 1 only
    ncut, ndof, nnz = run_gpu(ctx, mesh, phi)
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    assert ncut == 2 and ndof == O.number_p1if(om, phi)[1] == 5
    phi[:] = 1.0
    assert run_gpu(ctx, mesh, phi) == (0, 0, 0)


def test_levelset_size_checked(ctx):
    from drops_amd.mesh import KuhnMesh
    mesh = KuhnMesh.box(0.0, 1.0, 4)
    ctx.mesh(mesh)
    with pytest.raises(D.DhipError) as ei:
        ctx.levelset(np.zeros(10))
    assert ei.value.code == D.E_ARG


def test_full_size_c2_properties(ctx):
    """Configuration 2 size (128^3 cells): counts, symmetry, row sums of the mass
    matrix = patch-weighted hat integrals, and LB annihilating constants."""
    mesh, phi = cases.c1_sphere(128)
    ncut, ndof, nnz = run_gpu(ctx, mesh, phi, coef=(1.0, 1.0, 1.0), lc=(1.0, 1.0, mesh.h))
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    assert np.array_equal(ctx.cut_tets(), O.classify(om, phi))
    g = ctx.csr()
    rp, col = g["rowptr"].astype(np.int64), g["colind"]
    rows = np.repeat(np.arange(ndof), np.diff(rp))
    assert np.all(np.diff(col)[np.diff(rows) == 0] > 0)  # sorted, unique columns
    import scipy.sparse as sp
    for k in ("mass", "lb", "stab", "comb"):
        A = sp.csr_matrix((g[k], col, rp), shape=(ndof, ndof))
        assert abs(A - A.T).max() <= 1e-15 * abs(A).max(), k
    LB = sp.csr_matrix((g["lb"], col, rp), shape=(ndof, ndof))
    assert np.abs(LB @ np.ones(ndof)).max() < 1e-12 * abs(LB).max()
    M = sp.csr_matrix((g["mass"], col, rp), shape=(ndof, ndof))
    area = np.ones(ndof) @ (M @ np.ones(ndof))
    assert abs(area - 4 * np.pi) / (4 * np.pi) < 1e-3
