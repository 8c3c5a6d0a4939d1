"""Pins of the oracle's adaptive BDF1/BDF2 time loop (surfphasesep/surfphasesep.cpp:112-123,
396-491), no GPU: the BDF2 weights against exactness for polynomials in time, the step-size
function against its defining cases, the BDF2 system against the implicit-Euler system it reduces
to for equal history, the scheme's second order in time against implicit Euler's first, and the
loop's bookkeeping (first steps, retries, the carried error, the accepted solve's system)."""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.sparse as sp

import cases
import oracle as O

EPS = 0.05


def setup_torus(n):
    m, phi = cases.torus_case(n)
    om = O.make_mesh(m.orig, m.e, m.n)
    dof, nd = O.number_p1if(om, phi)
    r = O.assemble_p1(om, phi, dof, nd, coef=(1.0, 1.0, 1.0))
    return m, phi, om, dof, nd, r


@pytest.mark.parametrize("r", [0.25, 0.7, 1.0, 1.6, 3.0])
def test_bdf2_weights_exact_for_quadratics(r):
    """alpha q(t_{n+1}) - F_chi/M equals q'(t_{n+1}) for every quadratic q (the variable-step BDF2
    formula), and the extrapolation reproduces every linear q at t_{n+1}."""
    dt_prev = 0.013
    dt = r * dt_prev
    dte, w1, w2, w3, w4 = O.bdf2_weights(dt, dt_prev)
    t1, t0, tm = 0.4, 0.4 - dt, 0.4 - dt - dt_prev
    for a, b, c in [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (0.3, -2.0, 5.0)]:
        q = lambda t: a + b * t + c * t * t
        dq = b + 2 * c * t1
        cr = w3 * q(t0) + w4 * q(tm)
        assert abs((q(t1) - cr) / dte - dq) <= 1e-9 * (1 + abs(dq))
    for a, b in [(1.0, 0.0), (0.0, 1.0), (-0.4, 2.5)]:
        lin = lambda t: a + b * t
        assert abs(w1 * lin(t0) + w2 * lin(tm) - lin(t1)) <= 1e-13
    # a cubic is not differentiated exactly: the check above has teeth
    cub = lambda t: t ** 3
    assert abs((cub(t1) - (w3 * cub(t0) + w4 * cub(tm))) / dte - 3 * t1 ** 2) > 1e-6


def test_step_size_function():
    """F(0, dt) is the initial step whatever dt; at e = tol it is rho dt; it falls with e and is
    clamped to [min_step, max_step] (surfphasesep/surfphasesep.cpp:117-123)."""
    p = O.bdf_params(0.01, 0.005, rho=0.9, min_step=0.002, max_step=0.05)
    assert O.bdf_step_size(p, 0.0, 0.3) == 0.01
    assert O.bdf_step_size(p, 0.005, 0.02) == pytest.approx(0.018, rel=1e-15)
    assert O.bdf_step_size(p, 0.02, 0.02) == pytest.approx(0.009, rel=1e-15)  # sqrt(1/4) = 1/2
    assert O.bdf_step_size(p, 1e-12, 0.02) == 0.05
    assert O.bdf_step_size(p, 1e3, 0.02) == 0.002
    es = [1e-4, 1e-3, 5e-3, 2e-2]
    fs = [O.bdf_step_size(p, e, 0.01) for e in es]
    assert all(x > y for x, y in zip(fs, fs[1:]))


def test_bdf2_with_equal_history_is_implicit_euler_with_two_thirds_step():
    """c_{n-1} = c_n and r = 1: c_e = c_n, c_r = c_n and dt_eff = 2 dt / 3, so the BDF2 system is
    the implicit-Euler system of step 2 dt / 3 (rhs to rounding of c_r, all else bit for bit)."""
    m, phi, om, dof, nd, r = setup_torus(10)
    c = 0.5 + np.random.default_rng(7).uniform(-0.05, 0.05, nd)
    dt = 0.01
    dte, w1, w2, w3, w4 = O.bdf2_weights(dt, dt)
    assert dte == pytest.approx(2 * dt / 3, rel=1e-15)
    ce, cr = w1 * c + w2 * c, w3 * c + w4 * c
    args = (om, phi, dof, nd, r["rowptr"], r["colind"], r["mass"], r["lb"], r["stab"])
    b_bdf = O.ch_system_bdf(*args, ce, cr, dte, EPS, m.h)
    b_ie = O.ch_system(*args, c, dte, EPS, m.h)
    for a, b in zip(b_bdf[:3], b_ie[:3]):
        assert np.array_equal(a, b)
    assert np.array_equal(b_bdf[4], b_ie[4])
    assert np.array_equal(b_bdf[3][nd:], b_ie[3][nd:])
    assert np.abs(b_bdf[3][:nd] - b_ie[3][:nd]).max() <= 1e-15 * np.abs(b_ie[3][:nd]).max() * 4


def _run_fixed(setup, c0, T, N, bdf=True):
    m, phi, om, dof, nd, r = setup
    if bdf:
        p = O.bdf_params(T / N, 0.005, adaptive=False)
        s = O.BdfStepper(om, phi, dof, nd, r, c0, np.zeros(nd), p, EPS, m.h, tol=1e-13)
        for _ in range(N):
            st = s.step()
        assert st["t"] == pytest.approx(T, rel=1e-12)
        return s.c.copy()
    c, mu = c0.copy(), np.zeros(nd)
    for _ in range(N):
        o = O.ch_step(om, phi, dof, nd, r["rowptr"], r["colind"], r["mass"], r["lb"], r["stab"], c, mu,
                      dt=T / N, eps=EPS, rho=m.h, tol=1e-13)
        c, mu = o["c"], o["mu"]
    return c


def test_bdf2_is_second_order_in_time():
    """Fixed steps T/N: the BDF loop's error against a fine run falls like N^-2 (observed order
    >= 1.7 between N = 16 and 32), implicit Euler's like N^-1 (<= 1.2).  A dropped or wrong BDF2
    weight leaves first order."""
    setup = setup_torus(8)
    m, phi, om, dof, nd, r = setup
    M = sp.csr_matrix((r["mass"], r["colind"], r["rowptr"]), shape=(nd, nd))
    nrm = lambda v: math.sqrt(v @ (M @ v))
    c0 = 0.5 + np.random.default_rng(3).uniform(-0.05, 0.05, nd)
    T = 0.02
    cref = _run_fixed(setup, c0, T, 384)
    e = {N: nrm(_run_fixed(setup, c0, T, N) - cref) for N in (16, 32)}
    ei = {N: nrm(_run_fixed(setup, c0, T, N, bdf=False) - cref) for N in (16, 32)}
    order = math.log2(e[16] / e[32])
    order_ie = math.log2(ei[16] / ei[32])
    assert order >= 1.7, (e, order)
    assert order_ie <= 1.2, (ei, order_ie)
    assert e[32] < 0.2 * ei[32]


def test_adaptive_loop_bookkeeping():
    """The first step is BDF1 at step_size with e = 0; the second starts at step_size again
    (e is still 0); every accepted step has e <= tol or dt at a bound; a step taken at the first
    attempt starts from F(e, dt) of the step before; t is the sum of the accepted steps; some step
    is retried; and the accepted state solves the BDF2 system composed from the weights."""
    m, phi, om, dof, nd, r = setup_torus(12)
    c0 = 0.5 + np.random.default_rng(3).uniform(-0.05, 0.05, nd)
    p = O.bdf_params(0.01, 0.005, rho=0.9, adaptive=True, max_tries=30)
    s = O.BdfStepper(om, phi, dof, nd, r, c0, np.zeros(nd), p, EPS, m.h, tol=1e-11)
    stats, t = [], 0.0
    args = (om, phi, dof, nd, r["rowptr"], r["colind"], r["mass"], r["lb"], r["stab"])
    for k in range(8):
        cn, cp = s.c.copy(), s.cprev.copy()
        dt_prev = s.state.dt
        st = s.step()
        stats.append(st)
        t += st["dt"]
        assert st["step"] == k + 1 and st["t"] == pytest.approx(t, rel=1e-14)
        if k == 0:
            assert st["dt"] == 0.01 and st["e"] == 0.0 and st["iters_bdf2"] == 0 and st["tries"] == 1
            continue
        assert st["e"] <= p.tol or not (p.min_step < st["dt"] < p.max_step)
        if k == 1:
            assert st["dt"] == 0.01 or st["tries"] > 1
        elif st["tries"] == 1:
            assert st["dt"] == O.bdf_step_size(p, stats[k - 1]["e"], stats[k - 1]["dt"])
        dte, w1, w2, w3, w4 = O.bdf2_weights(st["dt"], dt_prev)
        brp, bcol, bval, b, _ = O.ch_system_bdf(*args, w1 * cn + w2 * cp, w3 * cn + w4 * cp, dte, EPS, m.h)
        K = sp.csr_matrix((bval, bcol, brp), shape=(2 * nd, 2 * nd))
        assert np.linalg.norm(b - K @ s.x) <= 2e-11 * np.linalg.norm(b)
        assert np.array_equal(s.cprev, cn)
    assert any(st["tries"] > 1 for st in stats)


def test_max_tries_leaves_state_unchanged():
    m, phi, om, dof, nd, r = setup_torus(10)
    c0 = 0.5 + np.random.default_rng(4).uniform(-0.05, 0.05, nd)
    p = O.bdf_params(0.01, 1e-12, adaptive=True, max_tries=2)
    s = O.BdfStepper(om, phi, dof, nd, r, c0, np.zeros(nd), p, EPS, m.h, tol=1e-10)
    s.step()
    x1, st1 = s.x.copy(), (s.state.step, s.state.t, s.state.dt, s.state.e)
    with pytest.raises(O.OracleError):
        s.step()
    assert np.array_equal(s.x, x1)
    assert (s.state.step, s.state.t, s.state.dt, s.state.e) == st1
