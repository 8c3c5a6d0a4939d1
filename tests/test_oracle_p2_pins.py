"""Pins of the oracle's P2 interface space (no GPU): the numbering against a first-encounter
restatement and a known DOF count, the P2 mass / Laplace-Beltrami / stabilisation forms
against closed-form integrals on planar interfaces (where P2 interpolation of quadratics
is exact and the composite Quad5 rule integrates the degree-4 integrands exactly), against
the P1 forms for linear functions on an oblique plane, and against the kernel and area
identities on curved surfaces."""
import numpy as np
import pytest
import scipy.sparse as sp

import cases
import oracle as O
from drops_amd.mesh import KuhnMesh, plane_x


def mesh_of(m):
    return O.make_mesh(m.orig, m.e, m.n)


def csr(r, key, n):
    return sp.csr_matrix((r[key], r["colind"], r["rowptr"]), shape=(n, n))


def plane_setup(n=10, off=0.33):
    m = KuhnMesh.box(0.0, 1.0, n)
    phi = m.sample_p2(plane_x(off))
    om = mesh_of(m)
    dof, nd = O.number_p2if(om, phi)
    r = O.assemble_p2(om, phi, dof, nd, rhs_kind=O.RHS_CONST, rhs_param=[1.0])
    return m, phi, om, dof, nd, r


def interp(m, dof, nd, fn):
    """dof vector of the P2 nodal interpolant of fn"""
    U = m.sample_p2(fn).ravel()
    u = np.zeros(nd)
    sel = dof >= 0
    u[dof[sel]] = U[sel]
    return u


def test_numbering_known_count_on_plane():
    """x = 0.33 on the unit cube, 10^3: every tet of the cell slab 0.3 < x < 0.4 is cut and no P2
    basis function vanishes on the plane, so the DOFs are exactly the 3 x 21 x 21 half-grid
    nodes of that slab."""
    m, phi, om, dof, nd, r = plane_setup()
    assert nd == 3 * 21 * 21
    x, _, _ = m.halfgrid_coords()
    xs = x.ravel()[dof >= 0]
    assert np.all((xs > 0.3 - 1e-12) & (xs < 0.4 + 1e-12))


@pytest.mark.parametrize("name", ["c1_sphere16", "torus24", "skew_brick"])
def test_numbering_first_encounter(name):
    """CreateNumbOnInterfaceP2: cut tets in order, local nodes 0..9 in order, a node gets the next
    number on its first encounter in a tet where its squared basis function has a positive
    interface integral."""
    m, phi = cases.PARITY_CASES[name]()
    om = mesh_of(m)
    dof, nd = O.number_p2if(om, phi)
    cut = O.classify(om, phi)
    nodes = cases.tet_p2_nodes(m, cut)
    ref = np.full(dof.size, -1, np.int32)
    k = 0
    for c, t in enumerate(cut):
        loc = O.tet_local_p2(om, phi, int(t))
        if loc.nfacet == 0:
            continue
        hs = np.ctypeslib.as_array(loc.hat_sq)
        for i in range(10):
            g = nodes[c, i]
            if ref[g] < 0 and hs[i] > 0.0:
                ref[g] = k
                k += 1
    assert k == nd
    assert np.array_equal(ref, dof)


@pytest.mark.parametrize("fn,mass,lb", [
    (lambda x, y, z: y, 1 / 3, 1.0),
    (lambda x, y, z: y * y, 1 / 5, 4 / 3),
    (lambda x, y, z: y * z, 1 / 9, 2 / 3),
    (lambda x, y, z: 1.0 + 0 * x, 1.0, 0.0),
    (lambda x, y, z: z * z - y + 0.5, 1 / 5 + 1 / 3 + 1 / 4 - 1 / 3 + 1 / 3 - 1 / 2, 4 / 3 + 1.0),
])
def test_plane_closed_forms(fn, mass, lb):
    """On x = 0.33 (area 1) the P2 interpolant of a quadratic in (y, z) is exact, so
    u'Mu = int u^2, u'Au = int |grad_G u|^2, and u'Su = 0 (u does not vary along the normal)."""
    m, phi, om, dof, nd, r = plane_setup()
    u = interp(m, dof, nd, fn)
    M, A, S = csr(r, "mass", nd), csr(r, "lb", nd), csr(r, "stab", nd)
    assert abs(u @ (M @ u) - mass) < 1e-13
    assert abs(u @ (A @ u) - lb) < 1e-12
    assert abs(u @ (S @ u)) < 1e-12 * max(1.0, np.abs(S).max())


def test_plane_stabilisation_and_load():
    """u = x: n.grad u = 1 on every cut tet, so u'Su is the volume of the cut slab (0.1); the
    constant load vector b_i = int_G phi_i gives b.u = int_G u."""
    m, phi, om, dof, nd, r = plane_setup()
    S = csr(r, "stab", nd)
    ux = interp(m, dof, nd, lambda x, y, z: x)
    assert abs(ux @ (S @ ux) - 0.1) < 1e-13
    uy = interp(m, dof, nd, lambda x, y, z: y * y + z)
    assert abs(r["rhs"] @ uy - (1 / 3 + 1 / 2)) < 1e-13
    assert abs(r["rhs"].sum() - 1.0) < 1e-13


def test_oblique_plane_matches_p1_for_linear_functions():
    """An oblique planar interface: for a linear u both spaces hold it exactly, the P2 normal
    equals the facet normal, and both rules integrate the (degree <= 2) integrands exactly.
    So the P2 and P1 quadratic forms agree; the P1 stabilisation carries the 1/6 factor of
    LocalNormalLaplaceBulkP1CL, the P2 one is the true volume integral."""
    m = KuhnMesh.box(0.0, 1.0, 8, shift=(1e-3, 2e-3, 3e-3))
    phi = m.sample_p2(lambda x, y, z: (x + 2 * y + 2 * z) / 3 - 0.8)
    om = mesh_of(m)
    d2, n2 = O.number_p2if(om, phi)
    r2 = O.assemble_p2(om, phi, d2, n2)
    d1, n1 = O.number_p1if(om, phi)
    r1 = O.assemble_p1(om, phi, d1, n1)
    for fn in (lambda x, y, z: x + y + z + 0.3, lambda x, y, z: 2 * x - y + 0.5 * z - 1.0):
        u2 = interp(m, d2, n2, fn)
        V = m.sample_p2(fn)[::2, ::2, ::2].ravel()
        u1 = np.zeros(n1)
        u1[d1[d1 >= 0]] = V[d1 >= 0]
        # the P1 form uses the QR facet normals of tiny cut facets, the P2 form the level-set
        # gradient: both exact up to rounding, with some cancellation in the projections
        for key, fac, tol in (("mass", 1.0, 1e-12), ("lb", 1.0, 1e-10), ("stab", 6.0, 1e-12)):
            q2 = u2 @ (csr(r2, key, n2) @ u2)
            q1 = u1 @ (csr(r1, key, n1) @ u1)
            assert abs(q2 - fac * q1) <= tol * abs(q2), key


@pytest.mark.parametrize("name", ["c1_sphere16", "skew_brick"])
def test_curved_identities(name):
    """Partition of unity: 1'M1 is the Quad5 area of the patch (the P1 oracle's area), and
    the Laplace-Beltrami and stabilisation forms annihilate constants.  Symmetric matrices,
    sorted columns."""
    m, phi = cases.PARITY_CASES[name]()
    om = mesh_of(m)
    dof, nd = O.number_p2if(om, phi)
    r = O.assemble_p2(om, phi, dof, nd)
    M, A, S = (csr(r, k, nd) for k in ("mass", "lb", "stab"))
    one = np.ones(nd)
    area = O.surface_area(om, phi, zero_policy=O.ZERO_THROW)
    assert abs(one @ (M @ one) - area) < 1e-12 * area
    assert np.abs(A @ one).max() < 1e-11 * np.abs(A).max()
    assert np.abs(S @ one).max() < 1e-11 * np.abs(S).max()
    for X in (M, A, S):
        assert abs(X - X.T).max() <= 1e-15 * abs(X).max()
    for i in range(nd):
        c = r["colind"][r["rowptr"][i]:r["rowptr"][i + 1]]
        assert np.all(np.diff(c) > 0)


def test_sphere_laplace_beltrami_converges():
    """int_G |grad_G x|^2 = (2/3) |G| on the unit sphere; the P2 interpolant of x on finer
    meshes approaches it (error well below the P1 one at the same mesh)."""
    errs = []
    for n in (8, 16):
        m, phi = cases.c1_sphere(n)
        om = mesh_of(m)
        dof, nd = O.number_p2if(om, phi)
        r = O.assemble_p2(om, phi, dof, nd)
        u = interp(m, dof, nd, lambda x, y, z: x)
        errs.append(abs(u @ (csr(r, "lb", nd) @ u) - 8 * np.pi / 3) / (8 * np.pi / 3))
    assert errs[1] < errs[0] and errs[1] < 0.02


def test_torus96_counts():
    """The P2 space of the torus (R = 1, r = 0.3) at 96^3 on [-2,2]^3: 47,048 cut tets, 95,552
    DOFs and 2,177,312 nonzeros, the figures of an independent count of the unique half-grid
    nodes and node pairs of the cut tets."""
    from drops_amd import ch_run
    m, phi = ch_run.torus_mesh(96)
    om = mesh_of(m)
    dof, nd = O.number_p2if(om, phi, zero_policy=O.ZERO_ALLOW)
    assert nd == 95552
    r = O.assemble_p2(om, phi, dof, nd, zero_policy=O.ZERO_ALLOW)
    assert r["colind"].size == 2177312


# ---------------------------------------------------------------- Cahn-Hilliard on the P2 space
def mob(x):
    return np.maximum((1.0 - x) * x, 0.0)


def test_p2_mobility_on_plane_closed_form():
    """x = 0.33: for u = y (|grad_G u| = 1) u' LM u = int_G m_h, m_h the P2 interpolant of the
    nodal mobilities, = sum_i Mob(c_i) int_G h_i = Mob(c) . b with b the constant-load vector.
    Also LM = Mob(c) LB for a constant concentration."""
    m, phi, om, dof, nd, r = plane_setup()
    rng = np.random.default_rng(5)
    c = rng.uniform(-0.1, 1.1, nd)
    lm, td, wl = O.ch_assemble(om, phi, dof, nd, r["rowptr"], r["colind"], c, space=2)
    LM = csr(dict(r, lm=lm), "lm", nd)
    u = interp(m, dof, nd, lambda x, y, z: y)
    assert abs(u @ (LM @ u) - mob(c) @ r["rhs"]) < 1e-13
    lm3, _, _ = O.ch_assemble(om, phi, dof, nd, r["rowptr"], r["colind"], np.full(nd, 0.3), space=2)
    assert np.abs(lm3 - mob(0.3) * r["lb"]).max() <= 1e-13 * np.abs(r["lb"]).max()


def test_p2_ch_vectors_and_symmetry():
    """M c and M f'(c) are the assembled P2 mass applied to c and f'(c); LM is symmetric and
    annihilates constants."""
    m, phi = cases.torus_case(12)
    om = mesh_of(m)
    dof, nd = O.number_p2if(om, phi)
    r = O.assemble_p2(om, phi, dof, nd)
    c = 0.5 + np.random.default_rng(2).uniform(-0.05, 0.05, nd)
    lm, td, wl = O.ch_assemble(om, phi, dof, nd, r["rowptr"], r["colind"], c, space=2)
    M = csr(r, "mass", nd)
    absM = abs(M)
    fp = (c - 1.0) * c * (c - 0.5)
    assert np.all(np.abs(td - M @ c) <= 1e-14 * (absM @ np.abs(c)))
    assert np.all(np.abs(wl - M @ fp) <= 1e-14 * (absM @ np.abs(fp)))
    LM = csr(dict(r, lm=lm), "lm", nd)
    assert np.abs(LM @ np.ones(nd)).max() <= 1e-11 * np.abs(lm).max()
    assert abs(LM - LM.T).max() <= 1e-15 * np.abs(lm).max()


def test_p2_ch_step_direct_solve_and_conservation():
    import scipy.sparse.linalg as spla
    m, phi = cases.torus_case(10)
    om = mesh_of(m)
    dof, nd = O.number_p2if(om, phi)
    r = O.assemble_p2(om, phi, dof, nd)
    rng = np.random.default_rng(3)
    c = 0.5 + rng.uniform(-0.05, 0.05, nd)
    mu = np.zeros(nd)
    M = csr(r, "mass", nd)
    total0 = np.ones(nd) @ (M @ c)
    for _ in range(2):
        o = O.ch_step(om, phi, dof, nd, r["rowptr"], r["colind"], r["mass"], r["lb"], r["stab"], c, mu,
                      dt=1e-4, eps=0.05, rho=m.h, tol=1e-10, space=2)
        assert o["ok"]
        K = sp.csr_matrix((o["val"], o["col"], o["rowptr"]), shape=(2 * nd, 2 * nd))
        xd = spla.spsolve(K.tocsc(), o["rhs"])
        x = np.concatenate([o["mu"], o["c"]])
        assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)
        c, mu = o["c"], o["mu"]
        assert abs(np.ones(nd) @ (M @ c) - total0) <= 1e-8 * abs(total0)
    assert 0.0 < c.min() and c.max() < 1.0
