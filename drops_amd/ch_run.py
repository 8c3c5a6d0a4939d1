"""Cahn-Hilliard trajectories on the trace-FEM surface (configuration: torus
phi = (sqrt(x^2+y^2) - 1)^2 + z^2 - 0.3^2 on [-2,2]^3, Kuhn brick of n^3 cells, implicit Euler
with dt = 1e-4, eps = 0.05, rho = h, sigma = S = 1 as CahnHilliardcGP1CL is built by
surfphasesep/separation.cpp:1128-1136; c_0 = 0.5 + U(-0.05, 0.05) per DOF, mu_0 = 0).

Everything runs through libdhip; the stepper keeps the state on the device."""
import time

import numpy as np

from . import dhip as D
from .mesh import KuhnMesh, config_shift, torus

DT, EPS, SIGMA, S_STAB = 1e-4, 0.05, 1.0, 1.0


def torus_mesh(n):
    m = KuhnMesh.box(-2.0, 2.0, n, shift=config_shift(4.0 / n))
    return m, m.sample_p2(torus(1.0, 0.3))


def initial_state(ndof, seed):
    rng = np.random.default_rng(seed)
    return 0.5 + rng.uniform(-0.05, 0.05, ndof), np.zeros(ndof)


def make_stepper(ctx, mesh, phi, restart=50, tol=1e-8, max_iter=20000, space=1):
    ctx.mesh(mesh)
    ctx.levelset(phi)
    return ctx.cahn_hilliard(DT, EPS, mesh.h, sigma=SIGMA, S=S_STAB, restart=restart, tol=tol, max_iter=max_iter,
                             space=space)


def run(ch, seed, steps, record=False, guess=None):
    """Run `steps` steps from the seeded initial state.  guess(history, system) -> start vector
    (device pointer or host array) or None for the previous-step start.  Returns per-step
    iterations and wall times, and (record=True) the states [mu; c] of every level."""
    c0, mu0 = initial_state(ch.n, seed)
    ch.set_state(c0, mu0)
    its = np.zeros(steps, np.int32)
    wall = np.zeros(steps)
    states = [np.concatenate([mu0, c0])] if record else None
    hist = [np.concatenate([mu0, c0])]
    for s in range(steps):
        t0 = time.perf_counter()
        ch.prepare()
        x0 = guess(hist, ch) if guess else None
        it, res, ok = ch.solve(x0=x0)
        wall[s] = time.perf_counter() - t0
        its[s] = it
        if record or guess:
            c, mu = ch.state()
            x = np.concatenate([mu, c])
            if record:
                states.append(x)
            hist.append(x)
            hist = hist[-4:]
    return its, wall, states
