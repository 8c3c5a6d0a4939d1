"""Pins of the learned-warm-start oracle (no GPU): the backward pass against central finite
differences of the exact (unrounded) network, the bfloat16 rounding against its definition,
and the features against their closed form."""
import numpy as np
import pytest

import mlp_oracle as M


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 2.0 ** -7, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0, 3.140625])
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6; representable values are unchanged
    assert list(M.bf16(x)) == [1.0, 1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -6, -2.5, 0.0, 3.140625]
    r = np.random.default_rng(0).normal(size=1000)
    assert np.all(np.abs(M.bf16(r) - r) <= 2.0 ** -8 * np.abs(r) * 1.0001)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    th = M.init_params(3)
    th += rng.normal(scale=0.05, size=th.size)
    X = rng.normal(size=(40, M.NF))
    T = rng.normal(scale=0.3, size=(40, M.NO))
    loss, g = M.loss_and_grad(th, X, T, rnd=M.identity)
    idx = rng.choice(th.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(th.size)
        e[i] = h
        lp = M.loss_and_grad(th + e, X, T, rnd=M.identity)[0]
        lm = M.loss_and_grad(th - e, X, T, rnd=M.identity)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_bf16_gradient_close_to_exact():
    """The straight-through gradient of the rounded network differs from the exact one by the
    rounding only (relative 1e-2 at these magnitudes)."""
    rng = np.random.default_rng(2)
    th = M.init_params(4)
    X = rng.normal(size=(200, M.NF))
    T = rng.normal(scale=0.3, size=(200, M.NO))
    _, g = M.loss_and_grad(th, X, T)
    _, ge = M.loss_and_grad(th, X, T, rnd=M.identity)
    assert np.linalg.norm(g - ge) <= 2e-2 * np.linalg.norm(ge)


def test_features_closed_form():
    n = 5
    rng = np.random.default_rng(3)
    hist = [rng.normal(size=2 * n) for _ in range(2)]
    res = rng.normal(size=2 * n)
    d = rng.uniform(1, 2, 2 * n)
    F, (sm, sc) = M.raw_features(hist, res, d, n)
    c0, c1 = hist[0][n:], hist[1][n:]
    assert sc == np.sqrt(np.mean((c0 - c1) ** 2)) or abs(sc / np.sqrt(np.mean((c0 - c1) ** 2)) - 1) < 1e-15
    assert np.array_equal(F[:, 0], c0) and np.allclose(F[:, 1], (c0 - c1) / sc, rtol=0, atol=0)
    assert np.sqrt(np.mean(F[:, 1] ** 2)) == pytest.approx(1.0, rel=1e-14)   # unit step scale
    assert not F[:, 2].any() and not F[:, 3].any()   # missing levels repeat the oldest
    assert np.array_equal(F[:, 9], (res[n:] / d[n:]) / sc) and np.array_equal(F[:, 11], d[n:])


def test_targets_invert_extrapolation():
    """A trajectory that is exactly linear in time has zero correction targets; the target
    of a general next state reproduces it through the extrapolation formula."""
    n = 7
    rng = np.random.default_rng(4)
    a, b = (rng.normal(size=2 * n) for _ in range(2))
    x = [a + b * t for t in range(4)]        # x_0 .. x_3
    hist = [x[2], x[1], x[0]]
    T = M.targets(x[3], hist, n, np.zeros(2), np.ones(2))
    assert np.abs(T).max() < 1e-10
    xn = rng.normal(size=2 * n)
    T = M.targets(xn, hist, n, np.zeros(2), np.ones(2))
    scale = M.step_scale(hist[0], hist[1], n)
    ex = M.extrapolation(hist, n, scale)
    back = np.concatenate([hist[0][:n] + scale[0] * (ex[:, 0] + T[:, 0]), hist[0][n:] + scale[1] * (ex[:, 1] + T[:, 1])])
    assert np.allclose(back, xn, rtol=1e-12, atol=1e-12)
