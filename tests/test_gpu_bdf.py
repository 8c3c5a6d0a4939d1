"""GPU parity of the adaptive BDF1/BDF2 loop (dhip_ch_bdf_step, surfphasesep/surfphasesep.cpp:
396-491) through the C ABI, on the P1 and P2 steppers.  Against the oracle's loop on the same
state: the same number of attempts per step, dt and t within 1e-8 relative, the error estimate
within 1e-6 relative (both come from GMRES solutions at tolerance 1e-11), the state within 1e-7
relative; every accepted device state solves the BDF2 system the oracle composes from the
device's own previous levels to 5e-11 relative residual.  Fixed steps, the restart of the loop by
set_state and by an implicit-Euler step, and the attempt bound are checked too."""
import numpy as np
import pytest
import scipy.sparse as sp

import cases
import oracle as O
from drops_amd import dhip as D

pytestmark = pytest.mark.gpu

EPS, TOL = 0.05, 1e-11


def problem(n, space):
    m, phi = cases.torus_case(n)
    om = O.make_mesh(m.orig, m.e, m.n)
    if space == 2:
        dof, nd = O.number_p2if(om, phi)
        r = O.assemble_p2(om, phi, dof, nd, coef=(1.0, 1.0, 1.0))
    else:
        dof, nd = O.number_p1if(om, phi)
        r = O.assemble_p1(om, phi, dof, nd, coef=(1.0, 1.0, 1.0))
    return m, phi, om, dof, nd, r


def make_ch(ctx, m, phi, space):
    ctx.mesh(m)
    ctx.levelset(phi)
    return ctx.cahn_hilliard(1e-4, EPS, m.h, space=space, tol=TOL, max_iter=20000)


@pytest.mark.parametrize("n,space", [(12, 1), (20, 1), (8, 2)])
def test_adaptive_bdf_matches_oracle(ctx, n, space):
    print(f"[gpu bdf] adaptive torus n={n} space={space}", flush=True)
    m, phi, om, dof, nd, r = problem(n, space)
    ch = make_ch(ctx, m, phi, space)
    assert ch.n == nd
    rng = np.random.default_rng(n)
    c0 = 0.5 + rng.uniform(-0.05, 0.05, nd)
    ch.set_state(c0, np.zeros(nd))
    prm = dict(step_size=0.01, tol=0.005, rho=0.9, min_step=0.0, max_step=float("inf"), adaptive=True,
               max_tries=30)
    o = O.BdfStepper(om, phi, dof, nd, r, c0, np.zeros(nd), O.bdf_params(**prm), EPS, m.h, tol=TOL,
                     space=space)
    args = (om, phi, dof, nd, r["rowptr"], r["colind"], r["mass"], r["lb"], r["stab"])
    retried = False
    cprev = None
    for k in range(8):
        cn, _ = ch.state()
        (g,) = ch.bdf_step(1, **prm)
        so = o.step()
        assert g["step"] == so["step"] == k + 1
        assert g["tries"] == so["tries"], (k, g, so)
        retried |= g["tries"] > 1
        assert g["dt"] == pytest.approx(so["dt"], rel=1e-8)
        assert g["t"] == pytest.approx(so["t"], rel=1e-8)
        assert g["e"] == pytest.approx(so["e"], rel=1e-6, abs=1e-300)
        c, mu = ch.state()
        xg = np.concatenate([mu, c])
        assert np.linalg.norm(xg - o.x) <= 1e-7 * np.linalg.norm(o.x)
        if k >= 1:
            # the device's accepted state solves the oracle's BDF2 system of the device's levels
            dte, w1, w2, w3, w4 = O.bdf2_weights(g["dt"], dt_prev)
            brp, bcol, bval, b, _ = O.ch_system_bdf(*args, w1 * cn + w2 * cprev, w3 * cn + w4 * cprev, dte, EPS,
                                                    m.h, space=space)
            K = sp.csr_matrix((bval, bcol, brp), shape=(2 * nd, 2 * nd))
            assert np.linalg.norm(b - K @ xg) <= 5e-11 * np.linalg.norm(b)
        cprev, dt_prev = cn, g["dt"]
    assert retried or space == 2
    ch.close()


def test_fixed_steps_and_restarts(ctx):
    print("[gpu bdf] fixed steps, restarts, attempt bound", flush=True)
    m, phi, om, dof, nd, r = problem(12, 1)
    ch = make_ch(ctx, m, phi, 1)
    c0 = 0.5 + np.random.default_rng(5).uniform(-0.05, 0.05, nd)
    ch.set_state(c0, np.zeros(nd))
    st = ch.bdf_step(4, step_size=0.004, adaptive=False)
    assert [s["step"] for s in st] == [1, 2, 3, 4]
    assert all(s["dt"] == 0.004 and s["tries"] == 1 for s in st)
    assert st[0]["iters_bdf2"] == 0 and all(s["iters_bdf2"] > 0 for s in st[1:])
    assert st[-1]["t"] == pytest.approx(0.016, rel=1e-14)
    o = O.BdfStepper(om, phi, dof, nd, r, c0, np.zeros(nd), O.bdf_params(0.004, 0.005, adaptive=False), EPS,
                     m.h, tol=TOL)
    for _ in range(4):
        o.step()
    c, mu = ch.state()
    assert np.linalg.norm(np.concatenate([mu, c]) - o.x) <= 1e-7 * np.linalg.norm(o.x)
    # set_state restarts the loop: step 1 again, t from 0
    ch.set_state(c0, np.zeros(nd))
    (s1,) = ch.bdf_step(1, step_size=0.004, adaptive=False)
    assert s1["step"] == 1 and s1["t"] == 0.004 and s1["iters_bdf2"] == 0
    # an implicit-Euler step makes the next BDF step a first step, time kept
    ch.step(1)
    (s2,) = ch.bdf_step(1, step_size=0.004, adaptive=False)
    assert s2["step"] == 1 and s2["iters_bdf2"] == 0 and s2["t"] == pytest.approx(0.008, rel=1e-14)
    # the attempt bound: the error cannot reach 1e-12, so the second step gives up, state kept
    c_before, mu_before = ch.state()
    with pytest.raises(D.DhipError) as ei:
        ch.bdf_step(1, step_size=0.004, tol=1e-12, adaptive=True, max_tries=2)
    assert ei.value.code == D.E_NOT_CONVERGED
    c_after, mu_after = ch.state()
    assert np.array_equal(c_before, c_after) and np.array_equal(mu_before, mu_after)
    ch.close()
