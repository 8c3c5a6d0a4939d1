"""GPU parity of the Cahn-Hilliard step (CahnHilliardcGP1CL, theta = 1, resting interface)
through the C ABI.  Bars: the step system bit-identical to the oracle except the mobility block
and the first preconditioner diagonal (1e-13 of their largest entry: the device contracts
precomputed geometric coefficients with the vertex mobilities); GMRES solutions with true relative
residual <= 2e-8 and within 1e-5 of the oracle's solution; several steps within 1e-6 of the
oracle's trajectory; the integral of c conserved to 1e-8."""
import numpy as np
import pytest
import scipy.sparse as sp

import cases
import oracle as O
from drops_amd import dhip as D

pytestmark = pytest.mark.gpu

DT, EPS = 1e-4, 0.05


def torus_problem(n):
    m, phi = cases.torus_case(n)
    om = O.make_mesh(m.orig, m.e, m.n)
    dof, nd = O.number_p1if(om, phi)
    r = O.assemble_p1(om, phi, dof, nd, coef=(1.0, 1.0, 1.0))
    return m, phi, om, dof, nd, r


def make_ch(ctx, m, phi, **kw):
    ctx.mesh(m)
    ctx.levelset(phi)
    return ctx.cahn_hilliard(DT, EPS, m.h, **kw)


@pytest.mark.parametrize("n", [12, 24])
def test_ch_system_matches_oracle(ctx, n):
    print(f"[gpu ch] system torus n={n}", flush=True)
    m, phi, om, dof, nd, r = torus_problem(n)
    ch = make_ch(ctx, m, phi)
    assert ch.n == nd
    rng = np.random.default_rng(n)
    c0 = 0.5 + rng.uniform(-0.05, 0.05, nd)
    mu0 = rng.normal(scale=1e-3, size=nd)
    ch.set_state(c0, mu0)
    ch.prepare()
    g = ch.system_host()
    o = O.ch_step(om, phi, dof, nd, r["rowptr"], r["colind"], r["mass"], r["lb"], r["stab"], c0, mu0,
                  dt=DT, eps=EPS, rho=m.h)
    assert np.array_equal(g["rowptr"], o["rowptr"].astype(np.int32))
    assert np.array_equal(g["col"], o["col"])
    # the mobility block: the device contracts precomputed geometric coefficients with the vertex
    # mobilities (the oracle evaluates the mobility at every quadrature point first) -> rounding
    # level; everything else bit-identical
    rp = o["rowptr"]
    in_a = np.zeros(o["val"].size, bool)
    for i in range(nd):
        ln = (rp[i + 1] - rp[i]) // 2
        in_a[rp[i]:rp[i] + ln] = True
    scale = np.abs(o["val"][in_a]).max()
    assert np.abs(g["val"][in_a] - o["val"][in_a]).max() <= 1e-13 * scale
    assert np.array_equal(g["val"][~in_a], o["val"][~in_a])
    assert np.array_equal(g["rhs"], o["rhs"])
    d3 = np.abs(o["pcdiag"][:nd]).max()
    assert np.abs(g["pcdiag"][:nd] - o["pcdiag"][:nd]).max() <= 1e-13 * d3
    assert np.array_equal(g["pcdiag"][nd:], o["pcdiag"][nd:])
    assert np.array_equal(g["state"], o["x0"])
    # the solve
    it, res, ok = ch.solve()
    assert ok and res <= 1e-8
    c1, mu1 = ch.state()
    K = sp.csr_matrix((o["val"], o["col"], o["rowptr"]), shape=(2 * nd, 2 * nd))
    x = np.concatenate([mu1, c1])
    assert np.linalg.norm(o["rhs"] - K @ x) <= 2e-8 * np.linalg.norm(o["rhs"])
    xo = np.concatenate([o["mu"], o["c"]])
    assert np.linalg.norm(x - xo) <= 1e-5 * np.linalg.norm(xo)
    assert abs(it - o["iters"]) <= max(5, 0.1 * o["iters"]), (it, o["iters"])
    ch.close()


def test_ch_trajectory_matches_oracle_and_conserves_mass(ctx):
    print("[gpu ch] trajectory torus n=16", flush=True)
    m, phi, om, dof, nd, r = torus_problem(16)
    ch = make_ch(ctx, m, phi)
    rng = np.random.default_rng(7)
    c = 0.5 + rng.uniform(-0.05, 0.05, nd)
    mu = np.zeros(nd)
    ch.set_state(c, mu)
    its, res = ch.step(5)
    assert np.all(res <= 1e-8) and np.all(its > 0)
    M = sp.csr_matrix((r["mass"], r["colind"], r["rowptr"]), shape=(nd, nd))
    total0 = np.ones(nd) @ (M @ c)
    for _ in range(5):
        o = O.ch_step(om, phi, dof, nd, r["rowptr"], r["colind"], r["mass"], r["lb"], r["stab"], c, mu,
                      dt=DT, eps=EPS, rho=m.h, tol=1e-12)
        c, mu = o["c"], o["mu"]
    cg, mug = ch.state()
    assert np.linalg.norm(cg - c) <= 1e-6 * np.linalg.norm(c)
    assert abs(np.ones(nd) @ (M @ cg) - total0) <= 1e-8 * abs(total0)
    ch.close()


def test_ch_start_vector_and_errors(ctx):
    print("[gpu ch] start vector / errors", flush=True)
    m, phi, om, dof, nd, r = torus_problem(12)
    ch = make_ch(ctx, m, phi)
    c0 = 0.5 + np.random.default_rng(9).uniform(-0.05, 0.05, nd)
    ch.set_state(c0, np.zeros(nd))
    with pytest.raises(D.DhipError) as ei:
        ch.solve()
    assert ei.value.code == D.E_STATE
    ch.prepare()
    it, res, ok = ch.solve()
    c1, mu1 = ch.state()
    # the same system again, started from its own solution: converged at iteration 0
    ch.set_state(c0, np.zeros(nd))
    ch.prepare()
    it2, res2, ok2 = ch.solve(x0=np.concatenate([mu1, c1]))
    assert ok2 and it2 == 0
    # a limit of 3 iterations reports non-convergence with the residual reached
    tight = ctx.cahn_hilliard(DT, EPS, m.h, max_iter=3)
    tight.set_state(c0, np.zeros(nd))
    tight.prepare()
    it3, res3, ok3 = tight.solve(raise_on_fail=False)
    assert not ok3 and it3 == 3 and res3 > 1e-8
    tight.close()
    with pytest.raises(D.DhipError) as ei:
        ctx.cahn_hilliard(-1.0, EPS, m.h)
    assert ei.value.code == D.E_ARG
    ch.close()


def test_gmres_diag_matches_oracle(ctx):
    print("[gpu ch] gmres_diag", flush=True)
    m, phi, om, dof, nd, r = torus_problem(12)
    c0 = 0.5 + np.random.default_rng(4).uniform(-0.05, 0.05, nd)
    o = O.ch_step(om, phi, dof, nd, r["rowptr"], r["colind"], r["mass"], r["lb"], r["stab"], c0, np.zeros(nd),
                  dt=DT, eps=EPS, rho=m.h)
    K = ctx.csr_create(o["rowptr"].astype(np.int32), o["col"], o["val"])
    for left in (False, True):
        x, it, res, ok = ctx.gmres_diag(K, o["pcdiag"], o["rhs"], x0=o["x0"], restart=50, tol=1e-8, left_pc=left)
        xo, ito, reso, oko = O.gmres_diag(o["rowptr"], o["col"], o["val"], o["pcdiag"], o["rhs"], o["x0"],
                                          restart=50, tol=1e-8, left_pc=left)
        assert ok and oko
        assert abs(it - ito) <= max(5, 0.1 * ito)
        assert np.linalg.norm(x - xo) <= 1e-5 * np.linalg.norm(xo)
    K.close()


def test_gmres_sell_xcd_mapping(ctx):
    """GMRES whose block slices exceed the LDS slice (the SELL-64 copy is read), 15 blocks of 384
    rows: XCD-aware row ranges (the default) against the plain block order.  Both converge with a
    true relative residual <= 2e-8, within 5% of each other's iteration count, and their solutions
    agree to the solve's accuracy."""
    m, phi, om, dof, nd, r = torus_problem(40)
    c0 = 0.5 + np.random.default_rng(4).uniform(-0.05, 0.05, nd)
    o = O.ch_step(om, phi, dof, nd, r["rowptr"], r["colind"], r["mass"], r["lb"], r["stab"], c0, np.zeros(nd),
                  dt=DT, eps=EPS, rho=m.h)
    n2 = o["rowptr"].size - 1
    assert (n2 + 383) // 384 == 15 and 385 * 4 + o["col"].size // 15 * 12 > 64 * 1024
    K = ctx.csr_create(o["rowptr"].astype(np.int32), o["col"], o["val"])
    Ks = sp.csr_matrix((o["val"], o["col"], o["rowptr"]), shape=(n2, n2))
    out = {}
    try:
        ctx.tune("pcg_rows_per_block", 384)
        for xm in (0, 1):
            ctx.tune("xcd_map", xm)
            print(f"[gpu ch] gmres sell xcd_map={xm}", flush=True)
            x, it, res, ok = ctx.gmres_diag(K, o["pcdiag"], o["rhs"], x0=o["x0"], restart=50, tol=1e-8)
            assert ok
            assert np.linalg.norm(o["rhs"] - Ks @ x) <= 2e-8 * np.linalg.norm(o["rhs"])
            out[xm] = (x, it)
    finally:
        ctx.tune("pcg_rows_per_block", 0)
        ctx.tune("xcd_map", 1)
    K.close()
    (x0, it0), (x1, it1) = out[0], out[1]
    assert abs(it0 - it1) <= max(5, 0.05 * it0)
    assert np.linalg.norm(x0 - x1) <= 1e-5 * np.linalg.norm(x0)
