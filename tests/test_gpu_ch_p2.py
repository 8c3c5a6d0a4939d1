"""GPU parity of the Cahn-Hilliard step on the P2 interface space through the C ABI.  Bars as
for the P1 stepper: the step system bit-identical to the oracle except the mobility block and
the first preconditioner diagonal (1e-13 of their largest entry: the device contracts
precomputed geometric coefficients with the nodal mobilities, the oracle evaluates the mobility
at every quadrature point first); solutions with true relative residual <= 2e-8 and within 1e-5
of the oracle's; a short trajectory within 1e-6 of the oracle's; the integral of c conserved."""
import numpy as np
import pytest
import scipy.sparse as sp

import cases
import oracle as O
from drops_amd import ch_run
from drops_amd import dhip as D

pytestmark = pytest.mark.gpu

DT, EPS = 1e-4, 0.05


def torus_problem_p2(n):
    m, phi = cases.torus_case(n)
    om = O.make_mesh(m.orig, m.e, m.n)
    dof, nd = O.number_p2if(om, phi)
    r = O.assemble_p2(om, phi, dof, nd, coef=(1.0, 1.0, 1.0))
    return m, phi, om, dof, nd, r


def make_ch(ctx, m, phi, **kw):
    ctx.mesh(m)
    ctx.levelset(phi)
    return ctx.cahn_hilliard(DT, EPS, m.h, space=2, **kw)


@pytest.mark.parametrize("n", [10, 16])
def test_ch_p2_system_matches_oracle(ctx, n):
    print(f"[gpu ch p2] system torus n={n}", flush=True)
    m, phi, om, dof, nd, r = torus_problem_p2(n)
    ch = make_ch(ctx, m, phi)
    assert ch.n == nd
    rng = np.random.default_rng(n)
    c0 = 0.5 + rng.uniform(-0.05, 0.05, nd)
    mu0 = rng.normal(scale=1e-3, size=nd)
    ch.set_state(c0, mu0)
    ch.prepare()
    g = ch.system_host()
    o = O.ch_step(om, phi, dof, nd, r["rowptr"], r["colind"], r["mass"], r["lb"], r["stab"], c0, mu0,
                  dt=DT, eps=EPS, rho=m.h, space=2)
    assert np.array_equal(g["rowptr"], o["rowptr"].astype(np.int32))
    assert np.array_equal(g["col"], o["col"])
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
    it, res, ok = ch.solve()
    assert ok and res <= 1e-8
    c1, mu1 = ch.state()
    K = sp.csr_matrix((o["val"], o["col"], o["rowptr"]), shape=(2 * nd, 2 * nd))
    x = np.concatenate([mu1, c1])
    assert np.linalg.norm(o["rhs"] - K @ x) <= 2e-8 * np.linalg.norm(o["rhs"])
    xo = np.concatenate([o["mu"], o["c"]])
    assert np.linalg.norm(x - xo) <= 1e-5 * np.linalg.norm(xo)
    ch.close()


def test_ch_p2_trajectory_matches_oracle_and_conserves_mass(ctx):
    print("[gpu ch p2] trajectory torus n=12", flush=True)
    m, phi, om, dof, nd, r = torus_problem_p2(12)
    ch = make_ch(ctx, m, phi)
    rng = np.random.default_rng(7)
    c = 0.5 + rng.uniform(-0.05, 0.05, nd)
    mu = np.zeros(nd)
    ch.set_state(c, mu)
    its, res = ch.step(3)
    assert np.all(res <= 1e-8) and np.all(its > 0)
    M = sp.csr_matrix((r["mass"], r["colind"], r["rowptr"]), shape=(nd, nd))
    total0 = np.ones(nd) @ (M @ c)
    for _ in range(3):
        o = O.ch_step(om, phi, dof, nd, r["rowptr"], r["colind"], r["mass"], r["lb"], r["stab"], c, mu,
                      dt=DT, eps=EPS, rho=m.h, tol=1e-12, space=2)
        c, mu = o["c"], o["mu"]
    cg, mug = ch.state()
    assert np.linalg.norm(cg - c) <= 1e-6 * np.linalg.norm(c)
    assert abs(np.ones(nd) @ (M @ cg) - total0) <= 1e-8 * abs(total0)
    ch.close()


def test_ch_p2_torus96_steps(ctx):
    """The full-size P2 system (torus 96^3, 2 x 95,552 unknowns): a few steps converge to the
    tolerance, c stays in (0, 1) and its integral is conserved."""
    print("[gpu ch p2] torus 96 steps", flush=True)
    m, phi = ch_run.torus_mesh(96)
    ch = ch_run.make_stepper(ctx, m, phi, space=2)
    assert ch.n == 95552
    c0, mu0 = ch_run.initial_state(ch.n, 1000)
    ch.set_state(c0, mu0)
    its, res = ch.step(3)
    print(f"[gpu ch p2] torus96 iterations {its.tolist()} resid {res.tolist()}", flush=True)
    assert np.all(res <= 1e-8)
    c, mu = ch.state()
    assert 0.0 < c.min() and c.max() < 1.0
    ch.close()
