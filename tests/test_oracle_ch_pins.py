"""Pins of the oracle's Cahn-Hilliard step pieces (no GPU): the mobility matrix against
closed forms on planar patches, the LocalMatVec vectors against the assembled mass matrix,
the block layout against scipy, and the step itself against a direct solve and the
conservation of the integral of c."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import cases
import oracle as O
from drops_amd.mesh import KuhnMesh, plane_x


def mesh_of(m):
    return O.make_mesh(m.orig, m.e, m.n)


def mob(x):
    return np.maximum((1.0 - x) * x, 0.0)


def plane_case():
    m = KuhnMesh.box(0.0, 1.0, 10)
    return m, m.sample_p2(plane_x(0.33))


def test_mobility_constant_concentration_is_scaled_laplace_beltrami():
    """Planar patch, no boundary facets, constant c: the Quad5 integral of the constant
    integrand m (P grad phi_i).(P grad phi_j) equals m times the facet sum of
    LocalLaplaceBeltramiP1CL."""
    m, phi = plane_case()
    om = mesh_of(m)
    cut = O.classify(om, phi)
    for t in cut[::7]:
        loc = O.tet_local(om, phi, t)
        if loc.nfacet == 0:
            continue
        lb = np.array(loc.lb).reshape(4, 4)
        lm = O.tet_mobility(om, phi, t, [0.3] * 4)
        assert np.abs(lm - mob(0.3) * lb).max() <= 1e-13 * np.abs(lb).max()


def test_mobility_linear_interpolant_on_plane():
    """Distinct vertex values: on a planar patch the projected gradients are constant, so
    LM_ij = (LB_ij / area) * sum_k Mob(c_k) int_Gamma lambda_k, and int_Gamma lambda_k is the
    row sum of the (exactly integrated) local mass matrix."""
    m, phi = plane_case()
    om = mesh_of(m)
    cut = O.classify(om, phi)
    rng = np.random.default_rng(1)
    checked = 0
    for t in cut[::5]:
        loc = O.tet_local(om, phi, t)
        if loc.nfacet == 0:
            continue
        cv = rng.uniform(-0.2, 1.2, 4)  # includes clipped (negative) mobilities
        lb = np.array(loc.lb).reshape(4, 4)
        mass = np.array(loc.mass).reshape(4, 4)
        lm = O.tet_mobility(om, phi, t, cv)
        expect = lb / loc.area * (mob(cv) @ mass.sum(axis=1))
        assert np.abs(lm - expect).max() <= 1e-12 * max(np.abs(lb).max(), 1e-300)
        checked += 1
    assert checked > 10


def setup_torus(n=12):
    m, phi = cases.torus_case(n)
    om = mesh_of(m)
    dof, nd = O.number_p1if(om, phi)
    r = O.assemble_p1(om, phi, dof, nd, coef=(1.0, 1.0, 1.0))
    return m, phi, om, dof, nd, r


def test_ch_vectors_are_mass_products():
    """LocalMatVecP1CL accumulations equal the assembled mass matrix applied to c and f'(c)."""
    m, phi, om, dof, nd, r = setup_torus()
    c = 0.5 + np.random.default_rng(2).uniform(-0.05, 0.05, nd)
    lm, td, wl = O.ch_assemble(om, phi, dof, nd, r["rowptr"], r["colind"], c)
    M = sp.csr_matrix((r["mass"], r["colind"], r["rowptr"]), shape=(nd, nd))
    absM = abs(M)
    fp = (c - 1.0) * c * (c - 0.5)
    assert np.all(np.abs(td - M @ c) <= 1e-14 * (absM @ np.abs(c)))
    assert np.all(np.abs(wl - M @ fp) <= 1e-14 * (absM @ np.abs(fp)))
    # constant concentration: LM = Mob(c) * (mobility-weighted Laplace-Beltrami) annihilates constants
    LMm = sp.csr_matrix((lm, r["colind"], r["rowptr"]), shape=(nd, nd))
    assert np.abs(LMm @ np.ones(nd)).max() <= 1e-12 * np.abs(lm).max()
    assert np.abs(LMm - LMm.T).max() <= 1e-15 * np.abs(lm).max()


def test_block_layout_matches_bmat():
    m, phi, om, dof, nd, r = setup_torus(8)
    rp, col = r["rowptr"], r["colind"]
    vals = [r["mass"], r["lb"], r["stab"], 2.0 * r["mass"]]
    brp, bcol, bval = O.block2x2(rp, col, *vals)
    K = sp.csr_matrix((bval, bcol, brp), shape=(2 * nd, 2 * nd))
    mats = [sp.csr_matrix((v, col, rp), shape=(nd, nd)) for v in vals]
    ref = sp.bmat([[mats[0], mats[1]], [mats[2], mats[3]]]).tocsr()
    assert abs(K - ref).max() == 0.0
    assert np.all(np.diff(bcol[brp[0]:brp[1]]) > 0)


def test_ch_step_direct_solve_and_conservation():
    m, phi, om, dof, nd, r = setup_torus()
    rng = np.random.default_rng(3)
    c = 0.5 + rng.uniform(-0.05, 0.05, nd)
    mu = np.zeros(nd)
    M = sp.csr_matrix((r["mass"], r["colind"], r["rowptr"]), shape=(nd, nd))
    total0 = np.ones(nd) @ (M @ c)
    for _ in range(3):
        o = O.ch_step(om, phi, dof, nd, r["rowptr"], r["colind"], r["mass"], r["lb"], r["stab"], c, mu,
                      dt=1e-4, eps=0.05, rho=m.h, tol=1e-10)
        assert o["ok"]
        K = sp.csr_matrix((o["val"], o["col"], o["rowptr"]), shape=(2 * nd, 2 * nd))
        xd = spla.spsolve(K.tocsc(), o["rhs"])
        x = np.concatenate([o["mu"], o["c"]])
        assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)
        c, mu = o["c"], o["mu"]
        assert abs(np.ones(nd) @ (M @ c) - total0) <= 1e-8 * abs(total0)
    # three short steps keep the concentration inside the two phases' range
    assert 0.0 < c.min() and c.max() < 1.0


def test_c_system_matches_numpy_composition():
    """orc_ch_system against the blocks composed here from the assembled matrices and the
    mobility matrix: LinComb orders as in CahnHilliardcGP1CL::DoStep/Update, bit for bit."""
    m, phi, om, dof, nd, r = setup_torus(10)
    c = 0.5 + np.random.default_rng(5).uniform(-0.05, 0.05, nd)
    dt, eps, rho, sigma, S = 1e-4, 0.05, m.h, 1.0, 1.0
    lm, td, wl = O.ch_assemble(om, phi, dof, nd, r["rowptr"], r["colind"], c)
    mass, lb, stab = r["mass"], r["lb"], r["stab"]
    A = (sigma * dt) * lm + (rho * dt) * stab
    B = 1.0 * (1.0 * mass + 0.0 * 0.0) + dt * 0.0
    Cb = 0.0 * lb + (-eps) * mass
    Dm = 1.0 * ((eps * eps) * lb + S * mass) + (rho * eps * eps) * stab
    brp, bcol, bval = O.block2x2(r["rowptr"], r["colind"], A, B, Cb, Dm)
    b2, c2, v2, rhs, pcd = O.ch_system(om, phi, dof, nd, r["rowptr"], r["colind"], mass, lb, stab, c, dt, eps, rho)
    assert np.array_equal(brp, b2) and np.array_equal(bcol, c2) and np.array_equal(bval, v2)
    assert np.array_equal(rhs, np.concatenate([1.0 * (td + dt * 0.0), ((-1.0) * wl + S * td) + 0.0]))
    rp = r["rowptr"]
    di = np.array([rp[i] + np.flatnonzero(r["colind"][rp[i]:rp[i + 1]] == i)[0] for i in range(nd)])
    P3 = 1.0 * (1.0 * mass + (sigma * dt) * lm) + (dt * rho) * stab
    P4 = 1.0 * (1.0 * mass + (eps * eps) * lb) + (eps * eps * rho) * stab
    assert np.array_equal(pcd, np.concatenate([P3[di], P4[di]]))
    # the full C step equals system + GMRES
    x = np.concatenate([np.zeros(nd), c])
    it, res, ok = O.ch_step_c(om, phi, dof, nd, r["rowptr"], r["colind"], mass, lb, stab, x, dt, eps, rho)
    xo, ito, reso, oko = O.gmres_diag(b2, c2, v2, pcd, rhs, np.concatenate([np.zeros(nd), c]), tol=1e-8)
    assert ok and oko and it == ito and np.array_equal(x, xo)
