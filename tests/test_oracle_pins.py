"""Pins of the CPU oracle (no GPU): the reference's known answers, brute-force
checks of the tables, closed forms for planar interfaces and exact triangle
integrals, so that a dropped term, a wrong index or a transposed operand in the
oracle fails here before GPU results are compared with it."""
import itertools

import numpy as np
import pytest

import cases
import oracle as O
from drops_amd.mesh import KuhnMesh, plane_x

EDGES = [(0, 1), (0, 2), (1, 2), (0, 3), (1, 3), (2, 3)]


def mesh_of(m):
    return O.make_mesh(m.orig, m.e, m.n)


# ---------------------------------------------------------------- known answers
def test_interfaceP1FE_known_answers():
    """tests/specifications/serial/TestinterfaceP1FE.ref: 8 (plane), 4 (plane through vertices,
    historical zero handling; the current reference throws), 1156 (sphere)."""
    m, p041, p040 = cases.planes()
    om = mesh_of(m)
    assert O.number_p1if(om, p041)[1] == 8
    with pytest.raises(O.OracleError) as ei:
        O.number_p1if(om, p040)
    assert ei.value.code == O.E_ZERO_NODE
    assert O.number_p1if(om, p040, zero_policy=O.ZERO_ALLOW)[1] == 4
    m, phi = cases.sphere20()
    assert O.number_p1if(mesh_of(m), phi)[1] == 1156


def test_single_tetra_known_answers():
    """(S. no int.) 0 and (S. int.) 4: a tet with one negative vertex numbers all 4 of its vertices."""
    m = KuhnMesh.box(0.0, 1.0, 1)
    om = mesh_of(m)
    phi = np.ones(m.halfgrid_shape)
    assert O.number_p1if(om, phi)[1] == 0
    phi[0, 0, 2] = -1.0  # vertex (1,0,0) of Kuhn tet 0
    loc = O.tet_local(om, phi, 0)
    assert loc.nfacet > 0 and all(loc.hat_sq[i] > 0 for i in range(4))


def test_principallattice_surface_known_answer():
    """Testprincipallattice.ref: surface 3.0909 (rel 1e-4) on [-1,1]^3, 8^3 cells, r = 0.5."""
    m, phi = cases.area_case()
    a = O.surface_area(mesh_of(m), phi)
    assert abs(a - 3.0909) / 3.0909 < 1e-4


# ---------------------------------------------------------------- tables
def test_lattice_tables_brute_force():
    b2, tets = O.lattice_tables()
    # 10 points of the order-2 lattice, barycentric coordinates in halves
    assert sorted(map(tuple, b2.tolist())) == sorted(
        p for p in itertools.product(range(3), repeat=4) if sum(p) == 2)
    pts = b2 / 2.0
    # the 8 sub-tets partition the reference tet: volumes 1/8 each, disjoint interiors
    vols = []
    for t in tets:
        X = pts[t][:, 1:]  # cartesian coordinates = barycentric 1..3
        vols.append(abs(np.linalg.det(X[1:] - X[0])) / 6)
    assert np.allclose(vols, 1 / 48) and abs(sum(vols) - 1 / 6) < 1e-15
    rng = np.random.default_rng(1)
    for _ in range(200):
        p = rng.dirichlet(np.ones(4))
        inside = 0
        for t in tets:
            lam = np.linalg.solve(np.vstack([pts[t].T[1:], np.ones(4)]), np.r_[p[1:], 1.0])
            inside += np.all(lam > -1e-12)
        assert inside == 1


def test_refpatch_table_brute_force():
    """For every sign pattern: the triangles tile the zero set of the linear function with
    those signs (areas against the convex hull of the roots), roots - 2 triangles,
    boundary flag exactly for 3 zero vertices."""
    size, tri, bnd = O.refpatch_table()
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    rng = np.random.default_rng(2)
    for s in itertools.product((-1, 0, 1), repeat=4):
        pid = s[0] + 3 * s[1] + 9 * s[2] + 27 * s[3] + 40
        if not any(s):
            continue
        zeros = [i for i in range(4) if s[i] == 0]
        cuts = [e for e, (a, b) in enumerate(EDGES) if s[a] * s[b] == -1]
        nroot = 0 if abs(sum(s)) == 4 else len(zeros) + len(cuts)
        assert size[pid] == (nroot - 2 if nroot >= 3 else 0)
        assert bnd[pid] == (len(zeros) == 3)
        if size[pid] == 0:
            continue
        vals = np.array([sv * rng.uniform(0.5, 2.0) for sv in s])
        pts = {}
        for i in zeros:
            pts[i] = V[i]
        for e in cuts:
            a, b = EDGES[e]
            al = vals[a] / (vals[a] - vals[b])
            pts[4 + e] = (1 - al) * V[a] + al * V[b]
        area = 0.0
        for k in range(size[pid]):
            P = [pts[r] for r in tri[pid, k]]
            area += np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0])) / 2
        allp = np.array(list(pts.values()))
        # area of the planar convex polygon through the roots
        c = allp.mean(0)
        nrm = np.cross(allp[1] - allp[0], allp[2] - allp[0])
        nrm /= np.linalg.norm(nrm)
        u = allp[0] - c
        u /= np.linalg.norm(u)
        w = np.cross(nrm, u)
        ang = np.arctan2((allp - c) @ w, (allp - c) @ u)
        poly = allp[np.argsort(ang)]
        ref = sum(np.linalg.norm(np.cross(poly[i] - c, poly[(i + 1) % len(poly)] - c)) / 2 for i in range(len(poly)))
        assert abs(area - ref) < 1e-12, s


# ---------------------------------------------------------------- patch geometry
@pytest.mark.parametrize("name", ["c1_sphere16", "torus24", "skew_brick"])
def test_patch_geometry_closed_forms(name):
    """QR absdet = 2 * triangle area, QR normal = +-unit cross product, cut vertices are
    zeros of the piecewise linear interpolant of the lattice values."""
    m, phi = cases.PARITY_CASES[name]()
    om = mesh_of(m)
    b2, tets = O.lattice_tables()
    cut = O.classify(om, phi)
    for t in cut[:: max(1, cut.size // 400)]:
        p = O.tet_patch(om, phi, int(t))
        assert p["nfacet"] > 0
        for f in range(p["nfacet"]):
            w = p["world"][f]
            cr = np.cross(w[1] - w[0], w[2] - w[0])
            a2 = np.linalg.norm(cr)
            assert abs(p["absdet"][f] - a2) <= 1e-13 * max(a2, 1e-30) + 1e-300
            if a2 > 1e-12 * m.h ** 2:
                assert abs(abs(p["normal"][f] @ cr) / a2 - 1) < 1e-12
            for j in range(3):
                bary = p["bary"][f, j]
                # linear interpolation of the level set on the lattice vertices around this point
                lam = bary * 2
                # the point lies on a lattice edge: its two lattice endpoints interpolate to zero
                ls = p["ls"]
                ends = [l for l in range(10) if np.all(np.abs(b2[l] / 2 - bary) <= 0.5 + 1e-12)]
                assert lam.sum() == pytest.approx(2.0)
                assert len(ends) >= 1
                vals = [ls[l] for l in ends]
                assert min(vals) <= 0 <= max(vals) or np.any(np.isclose(vals, 0))


def _exact_p1_integrals(p):
    """Exact integrals over the patch triangles of phi_i phi_j (degree 2) with the
    closed form int_T f g = A/12 (sum f_k g_k + sum f_k sum g_k) for linear f, g."""
    M = np.zeros((4, 4))
    for f in range(p["nfacet"]):
        A = np.linalg.norm(np.cross(p["world"][f][1] - p["world"][f][0], p["world"][f][2] - p["world"][f][0])) / 2
        A *= 0.5 if p["bnd"][f] else 1.0
        B = p["bary"][f]  # (3,4): phi_i at the triangle vertices
        M += A / 12 * (B.T @ B + np.outer(B.sum(0), B.sum(0)))
    return M


@pytest.mark.parametrize("name", ["c1_sphere16", "ellipsoid_ragged", "skew_brick"])
def test_mass_matches_exact_integrals(name):
    """Second route for LocalInterfaceMassP1CL (cf. tests/interfaceP1FE.cpp:140-230): composite
    Quad2 vs exact triangle integrals; and the alpha factor applies to off-diagonals only."""
    m, phi = cases.PARITY_CASES[name]()
    om = mesh_of(m)
    cut = O.classify(om, phi)
    for t in cut[:: max(1, cut.size // 300)]:
        p = O.tet_patch(om, phi, int(t))
        ref = _exact_p1_integrals(p)
        loc = O.tet_local(om, phi, int(t), coef=(0.5, 1.0, 1.0))
        M = np.ctypeslib.as_array(loc.mass).reshape(4, 4)
        alpha = np.where(np.eye(4, dtype=bool), 1.0, 0.5)
        assert np.allclose(M, alpha * ref, rtol=1e-12, atol=1e-15 * np.abs(ref).max())


def test_planar_closed_forms():
    """Plane x = 0.33 through the unit cube (10^3 cells): area 1, u=y gives u'Au = area and
    u'Mu = int y^2 = 1/3, u=x gives u'Cu = |layer|/6 (the /6 of LocalNormalLaplaceBulkP1CL),
    u=y gives u'Cu = 0, rhs of f=1 sums to the area."""
    m = KuhnMesh.box(0.0, 1.0, 10)
    phi = m.sample_p2(plane_x(0.33))
    om = mesh_of(m)
    dof, nd = O.number_p1if(om, phi)
    r = O.assemble_p1(om, phi, dof, nd, rhs_kind=O.RHS_CONST, rhs_param=[1.0])
    V = m.vertex_coords().reshape(-1, 3)
    vod = np.empty(nd, np.int64)
    vod[dof[dof >= 0]] = np.nonzero(dof >= 0)[0]
    X = V[vod]
    import scipy.sparse as sp
    mk = lambda k: sp.csr_matrix((r[k], r["colind"], r["rowptr"]), shape=(nd, nd))
    M, A, C = mk("mass"), mk("lb"), mk("stab")
    one = np.ones(nd)
    assert abs(one @ M @ one - 1.0) < 1e-13
    assert abs(r["rhs"].sum() - 1.0) < 1e-13
    y, x = X[:, 1], X[:, 0]
    assert abs(y @ A @ y - 1.0) < 1e-12
    assert abs(y @ M @ y - 1.0 / 3.0) < 1e-13
    assert abs(x @ C @ x - 0.1 / 6.0) < 1e-13
    assert abs(y @ C @ y) < 1e-13
    assert np.abs(A @ one).max() < 1e-12


def test_sphere_solution_converges():
    """Configuration 1 problem (-Delta_G + 1 + h-scaled stabilisation) u = 3xy - x^3 on the unit sphere:
    the discrete solution approaches u* = 3/7 xy - (x^3 - 3/5 x)/13 - x/5 (extended constantly along
    normals) at O(h^2) in L2 of the discrete surface: error ratio ~4 per mesh halving."""
    import scipy.sparse as sp
    errs = []
    for n in (16, 32):
        m, phi = cases.c1_sphere(n)
        om = mesh_of(m)
        dof, nd = O.number_p1if(om, phi)
        r = O.assemble_p1(om, phi, dof, nd, rhs_kind=O.RHS_3XY_MINUS_X3)
        comb = (1.0 * (r["mass"] + r["lb"])) + m.h * r["stab"]
        u, it, res, ok = O.pcg_jacobi(r["rowptr"], r["colind"], comb, r["rhs"], tol=1e-12)
        assert ok
        V = m.vertex_coords().reshape(-1, 3)
        vod = np.empty(nd, np.int64)
        vod[dof[dof >= 0]] = np.nonzero(dof >= 0)[0]
        P = V[vod] / np.linalg.norm(V[vod], axis=1)[:, None]
        x, y = P[:, 0], P[:, 1]
        ue = 3 / 7 * x * y - (x ** 3 - 0.6 * x) / 13 - x / 5
        M = sp.csr_matrix((r["mass"], r["colind"], r["rowptr"]), shape=(nd, nd))
        e = u - ue
        errs.append(np.sqrt(e @ M @ e) / np.sqrt(ue @ M @ ue))
    assert 3.6 < errs[0] / errs[1] < 4.5 and errs[1] < 4e-3, errs


# ---------------------------------------------------------------- linear algebra
def _random_csr(n, density, seed, spd=True):
    import scipy.sparse as sp
    M = sp.random(n, n, density=density, random_state=seed)
    M = (M + M.T) if spd else M
    M = (M + sp.eye(n) * (4.0 if spd else 6.0)).tocsr()
    M.sort_indices()
    return M


def test_spmv_and_kahan_dot():
    M = _random_csr(200, 0.05, 1)
    x = np.random.default_rng(0).normal(size=200)
    y = O.spmv(M.indptr.astype(np.int64), M.indices.astype(np.int32), M.data, x)
    assert np.allclose(y, M @ x, rtol=1e-14, atol=1e-14)
    import math
    a = np.random.default_rng(1).normal(size=5000) * 1e3
    b = np.random.default_rng(2).normal(size=5000)
    exact = math.fsum(a * b)  # products are exact-rounded in both
    assert abs(O.kahan_dot(a, b) - exact) <= 2 * np.finfo(float).eps * abs(exact) + 1e-12


def test_pcg_and_gmres_against_direct_solve():
    M = _random_csr(150, 0.04, 3)
    b = np.random.default_rng(4).normal(size=150)
    xd = np.linalg.solve(M.toarray(), b)
    rp, col = M.indptr.astype(np.int64), M.indices.astype(np.int32)
    x, it, res, ok = O.pcg_jacobi(rp, col, M.data, b, tol=1e-12)
    assert ok and np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-10
    N = _random_csr(150, 0.04, 5, spd=False)
    xd = np.linalg.solve(N.toarray(), b)
    rp, col = N.indptr.astype(np.int64), N.indices.astype(np.int32)
    for left in (False, True):
        x, it, res, ok = O.gmres_jacobi(rp, col, N.data, b, restart=20, tol=1e-12, left_pc=left)
        assert ok and np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9
    # the reference's GMRES(m) performs m-1 Arnoldi steps per cycle: with m=2 it is a
    # restarted steepest-descent-like method that still converges on this system
    x, it, res, ok = O.gmres_jacobi(rp, col, N.data, b, restart=2, tol=1e-10, max_iter=5000)
    assert ok


def _diag_k_distinct(k, n, seed):
    rng = np.random.default_rng(seed)
    d = np.repeat(np.arange(1, k + 1, dtype=float), n // k)
    rng.shuffle(d)
    return np.arange(n + 1, dtype=np.int64), np.arange(n, dtype=np.int32), d, rng.normal(size=n)


@pytest.mark.parametrize("left", [False, True])
def test_gmres_restart_m_takes_m_minus_1_steps(left):
    """The reference's GMRES(m) runs m-1 Arnoldi steps per cycle (num/krylovsolver.h:320-463).
    A diagonal matrix with k distinct eigenvalues has a minimal polynomial of degree k, so the
    exact solution lies in the k-th Krylov space and no earlier one (generic b): with restart
    k+1 the first cycle reaches it after exactly k steps; with restart k a cycle has only k-1
    steps and a second cycle is needed.  A GMRES(m) with m steps per cycle fails the second
    assertion, one with m-2 steps the first."""
    for k in (3, 5, 8):
        rp, col, d, b = _diag_k_distinct(k, 8 * k, k)
        x, it, res, ok = O.gmres_jacobi(rp, col, d, b, restart=k + 1, tol=1e-10, max_iter=500, left_pc=left,
                                        use_pc=False)
        assert ok and it == k, (k, it)
        assert np.abs(x - b / d).max() < 1e-12
        x, it, res, ok = O.gmres_jacobi(rp, col, d, b, restart=k, tol=1e-10, max_iter=500, left_pc=left,
                                        use_pc=False)
        assert ok and it > k, (k, it)


def test_pcg_failure_is_reported():
    M = _random_csr(100, 0.05, 7)
    b = np.ones(100)
    x, it, res, ok = O.pcg_jacobi(M.indptr.astype(np.int64), M.indices.astype(np.int32), M.data, b,
                                  max_iter=2, tol=1e-14)
    assert not ok and it == 2 and res > 1e-14
