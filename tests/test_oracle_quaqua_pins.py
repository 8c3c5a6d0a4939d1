"""Pins of the oracle's QuaQua-mapped P2 geometry (InterfaceCommonDataP2CL + QuaQuaMapperCL,
surfactant/ifacetransp.h:200-276, levelset/levelsetmapper.cpp:551-638, LocalMassP2CL /
LocalLaplaceBeltramiP2CL / LocalVectorP2CL ifacetransp.h:2302-2403) by what the mathematics fixes:
  - |x|^2 - 1 is quadratic, so its P2 interpolant is exact: the recovered gradient is 2x at every node,
    the base points lie on the unit sphere, the quasi-normal there is radial, and the mapped surface is
    the sphere itself: the mass matrix sums to 4 pi up to quadrature error (the planar patch misses by O(h^2));
  - the Jacobian of the base-point map equals its central differences (sphere and torus);
  - W^-1 turns the planar gradient of a linear function into the surface gradient of its lift, which the
    chain rule fixes from the Jacobian and the surface normal alone;
  - Laplace-Beltrami: rows sum to zero, symmetric, positive semi-definite, and the coordinate functions'
    Rayleigh quotient tends to 2 (the eigenvalue of the unit sphere);
  - the load vector of f = 1 is the mass matrix's row sums."""
import numpy as np
import pytest
import scipy.sparse as sp

import cases
import oracle as O


def mapped(n, case=cases.c1_sphere):
    mesh, phi = case(n)
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    return mesh, om, phi, O.recovered_gradient(om, phi)


def tet_vertices(mesh, t):
    n = mesh.n
    cell, k = t // 6, t % 6
    i = np.array([cell % n[0], (cell // n[0]) % n[1], cell // (n[0] * n[1])])
    idx = i[None, :] + cases.KUHN[k]
    return mesh.vertex_coords()[idx[:, 2], idx[:, 1], idx[:, 0]]   # (4, 3)


def node_coords(mesh):
    return np.stack([a.ravel() for a in mesh.halfgrid_coords()], axis=1)


def test_recovered_gradient_is_exact_for_a_quadratic_level_set():
    mesh, om, phi, g = mapped(12)
    X = node_coords(mesh)
    assert np.abs(g - 2.0 * X).max() <= 1e-12


def test_mapped_sphere_area_is_four_pi():
    errs = []
    for n in (16, 32):
        mesh, om, phi, g = mapped(n)
        dof, nd = O.number_p2if(om, phi)
        qq = O.assemble_p2qq(om, phi, g, dof, nd)
        pl = O.assemble_p2(om, phi, dof, nd)
        errs.append(abs(qq["mass"].sum() - 4 * np.pi))
        # the planar patch: O(h^2)
        assert abs(pl["mass"].sum() - 4 * np.pi) > 1e-2
    assert errs[0] < 3e-6 and errs[1] < 5e-8 and errs[0] / errs[1] > 30, errs


def test_base_points_on_the_sphere_along_the_radius():
    mesh, om, phi, g = mapped(16)
    cut = O.classify(om, phi)
    rng = np.random.default_rng(1)
    for t in rng.choice(cut, 40, replace=False):
        X = tet_vertices(mesh, int(t))
        b = rng.dirichlet(np.ones(4))
        x0 = X.T @ b
        bxb, y, dh, dph, it = O.quaqua_point(om, phi, g, int(t), b)
        assert abs(np.linalg.norm(y) - 1.0) < 1e-7
        assert np.abs(X.T @ bxb - y).max() < 1e-13          # base point given in the tet's coordinates
        assert np.linalg.norm(np.cross(x0 - y, y)) < 1e-12   # quasi-normal 2y: radial
        assert abs(np.linalg.norm(x0 - y) - abs(dh)) < 1e-7      # to the Newton tolerance on |F|
        assert 1 <= it <= 20


@pytest.mark.parametrize("case,n", [(cases.c1_sphere, 16), (cases.torus_case, 24)])
def test_jacobian_equals_central_differences(case, n):
    mesh, om, phi, g = mapped(n, case)
    cut = O.classify(om, phi)
    P = O.quaqua_params(maxiter=60, tol=1e-14)
    rng = np.random.default_rng(2)
    for t in rng.choice(cut, 12, replace=False):
        X = tet_vertices(mesh, int(t))
        A = np.vstack([X.T, np.ones(4)])
        b = rng.dirichlet(np.ones(4))
        _, _, _, dph, _ = O.quaqua_point(om, phi, g, int(t), b, P)
        eps = 1e-5 * mesh.h
        J = np.zeros((3, 3))
        for c in range(3):
            db = np.linalg.solve(A, np.r_[eps * np.eye(3)[c], 0.0])
            _, yp, _, _, _ = O.quaqua_point(om, phi, g, int(t), b + db, P)
            _, ym, _, _, _ = O.quaqua_point(om, phi, g, int(t), b - db, P)
            J[:, c] = (yp - ym) / (2 * eps)
        assert np.abs(J - dph).max() <= 1e-7 * np.abs(dph).max()


@pytest.mark.parametrize("case,n", [(cases.c1_sphere, 16), (cases.torus_case, 24)])
def test_winv_is_the_chain_rule(case, n):
    mesh, om, phi, g = mapped(n, case)
    cut = O.classify(om, phi)
    rng = np.random.default_rng(3)
    for t in rng.choice(cut, 12, replace=False):
        b = rng.dirichlet(np.ones(4))
        nl = rng.normal(size=3)
        nl /= np.linalg.norm(nl)
        Wi = O.quaqua_winv(om, phi, g, int(t), b, nl)
        _, y, _, dph, _ = O.quaqua_point(om, phi, g, int(t), b)
        # the surface normal at the base point: radial on the sphere, else the left null vector of dph
        ny = y / np.linalg.norm(y) if case is cases.c1_sphere else np.linalg.svd(dph)[0][:, 2]
        t1 = np.cross(nl, [1.0, 0.0, 0.0])
        t1 /= np.linalg.norm(t1)
        t2 = np.cross(nl, t1)
        a = rng.normal(size=3)   # u(x) = a.x on the plane with normal nl
        G = np.linalg.solve(np.vstack([dph @ t1, dph @ t2, ny]), [a @ t1, a @ t2, 0.0])
        assert np.abs(Wi @ (a - (a @ nl) * nl) - G).max() <= 1e-10 * np.abs(G).max()


def test_laplace_beltrami_and_load_vector():
    errs = []
    for n in (16, 32):
        mesh, om, phi, g = mapped(n)
        dof, nd = O.number_p2if(om, phi)
        r = O.assemble_p2qq(om, phi, g, dof, nd, rhs_kind=O.RHS_CONST, rhs_param=[1.0])
        M = sp.csr_matrix((r["mass"], r["colind"], r["rowptr"]), shape=(nd, nd))
        A = sp.csr_matrix((r["lb"], r["colind"], r["rowptr"]), shape=(nd, nd))
        assert abs(A - A.T).max() <= 1e-15 * abs(A).max() and abs(M - M.T).max() == 0.0
        assert np.abs(A @ np.ones(nd)).max() <= 1e-13 * abs(A).max()
        assert np.abs(r["rhs"] - M @ np.ones(nd)).max() <= 1e-15
        if n == 16:
            assert np.linalg.eigvalsh(A.toarray()).min() >= -1e-12 * abs(A).max()
        X = np.zeros((nd, 3))
        X[dof[dof >= 0]] = node_coords(mesh)[dof >= 0]
        R = [(X[:, c] @ (A @ X[:, c])) / (X[:, c] @ (M @ X[:, c])) for c in range(3)]
        errs.append(max(abs(q - 2.0) for q in R))
    assert errs[0] < 2e-2 and errs[1] < errs[0] / 3.0, errs


def test_mapped_pattern_is_the_lattice1_tets():
    """Tets cut only on lattice 2 (the vertex values of one sign) contribute to the planar P2 matrices but
    not to the mapped ones; every mapped entry is a planar one."""
    mesh, om, phi, g = mapped(12)
    dof, nd = O.number_p2if(om, phi)
    qq = O.assemble_p2qq(om, phi, g, dof, nd)
    pl = O.assemble_p2(om, phi, dof, nd)
    assert qq["colind"].size < pl["colind"].size
    Q = sp.csr_matrix((np.ones(qq["colind"].size), qq["colind"], qq["rowptr"]), shape=(nd, nd))
    Pm = sp.csr_matrix((np.ones(pl["colind"].size), pl["colind"], pl["rowptr"]), shape=(nd, nd))
    assert (Q.multiply(Pm)).nnz == Q.nnz
