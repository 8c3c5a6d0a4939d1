"""GPU parity of the P2 interface space against the CPU oracle, through the C ABI.

Bar: classification, numbering and CSR pattern bit-exact; local 10x10 forms, assembled
entries and load vector bit-exact (both sides evaluate every product and sum in the same
order without FMA contraction).  At the torus 96^3 size: the DOF and nonzero counts of
the independent count, symmetry, the kernel of the Laplace-Beltrami and stabilisation
operators, and 1'M1 = the patch area."""
import numpy as np
import pytest

import cases
import oracle as O
from drops_amd import dhip as D

pytestmark = pytest.mark.gpu

COEF, LC = (1.0, 0.7, 1.3), (1.0, 1.0, 0.1)


def run_p2(ctx, mesh, phi, zero=D.ZERO_THROW, rhs=D.RHS_3XY_MINUS_X3):
    print(f"[gpu p2] mesh n={mesh.n}", flush=True)
    ctx.mesh(mesh)
    ctx.levelset(phi)
    return ctx.assemble_p2(zero_policy=zero, coef=COEF, rhs_kind=rhs, lc=LC)


@pytest.mark.parametrize("name", sorted(cases.PARITY_CASES))
def test_p2_assembly_bit_exact(ctx, name):
    mesh, phi = cases.PARITY_CASES[name]()
    ncut, ndof, nnz = run_p2(ctx, mesh, phi)
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    cut = ctx.cut_tets()
    assert np.array_equal(cut, O.classify(om, phi))
    loc, flags = ctx.local_p2()
    for c in range(ncut):
        r = O.tet_local_p2(om, phi, int(cut[c]), coef=COEF, rhs_kind=O.RHS_3XY_MINUS_X3)
        ref = np.concatenate([np.ctypeslib.as_array(r.mass), np.ctypeslib.as_array(r.lb),
                              np.ctypeslib.as_array(r.stab), np.ctypeslib.as_array(r.rhs),
                              np.ctypeslib.as_array(r.hat_sq)])
        assert bool(flags[c] & 1) == (r.nfacet > 0)
        if not np.array_equal(loc[c], ref):
            d = np.abs(loc[c] - ref) / np.maximum(np.abs(ref), 1e-300)
            raise AssertionError(f"tet {cut[c]}: P2 local data differ, max rel {d.max():.3e} at {int(d.argmax())}")
    dof_ref, nd_ref = O.number_p2if(om, phi)
    assert ndof == nd_ref
    nodes = cases.tet_p2_nodes(mesh, cut)
    dofs = ctx.tet_dofs()
    assert np.array_equal(dofs, dof_ref[nodes])
    assert np.array_equal(dof_ref[ctx.dof_nodes()], np.arange(ndof))
    g = ctx.csr()
    ref = O.assemble_p2(om, phi, dof_ref, nd_ref, coef=COEF, rhs_kind=O.RHS_3XY_MINUS_X3)
    assert nnz == ref["colind"].size
    assert np.array_equal(g["rowptr"].astype(np.int64), ref["rowptr"])
    assert np.array_equal(g["colind"], ref["colind"])
    for k in ("mass", "lb", "stab", "rhs"):
        assert np.array_equal(g[k], ref[k]), k
    comb = (1.0 * (LC[0] * ref["mass"] + LC[1] * ref["lb"])) + LC[2] * ref["stab"]
    assert np.array_equal(g["comb"], comb)


def test_p2_torus96_counts_and_identities(ctx):
    """The P2 space of the torus at 96^3: 95,552 DOFs and 2,177,312 nonzeros; the operators
    are symmetric, LB and stabilisation annihilate constants, 1'M1 is the patch area."""
    from drops_amd import ch_run
    import scipy.sparse as sp
    mesh, phi = ch_run.torus_mesh(96)
    ncut, ndof, nnz = run_p2(ctx, mesh, phi, zero=D.ZERO_ALLOW, rhs=D.RHS_NONE)
    assert (ncut, ndof, nnz) == (47048, 95552, 2177312)
    g = ctx.csr()
    mats = {k: sp.csr_matrix((g[k], g["colind"], g["rowptr"]), shape=(ndof, ndof)) for k in ("mass", "lb", "stab")}
    one = np.ones(ndof)
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    area = O.surface_area(om, phi, zero_policy=O.ZERO_ALLOW)
    assert abs(one @ (mats["mass"] @ one) - area) < 1e-12 * area
    for k in ("lb", "stab"):
        X = mats[k]
        assert np.abs(X @ one).max() < 1e-10 * abs(X).max(), k
    for X in mats.values():
        assert abs(X - X.T).max() <= 1e-15 * abs(X).max()


def test_p2_single_cell_and_empty(ctx):
    """One brick cell with one negative corner, and an interface-free level set."""
    from drops_amd.mesh import KuhnMesh
    mesh = KuhnMesh.box(0.0, 1.0, 1)
    phi = np.ones(mesh.halfgrid_shape)
    phi[0, 0, 2] = -1.0
    ncut, ndof, nnz = run_p2(ctx, mesh, phi)
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    assert ncut == 2 and ndof == O.number_p2if(om, phi)[1]
    phi[:] = 1.0
    assert run_p2(ctx, mesh, phi) == (0, 0, 0)
    assert ctx.csr()["rowptr"].tolist() == [0]
