"""CPU oracle of the learned warm start (test infrastructure only).

Float64 restatement of the per-DOF network the library trains and evaluates on the GPU:
inputs of every layer and all weights are rounded to bfloat16 (round to nearest even) as the
device's matrix-core products take them, products and sums are float64 here (float32 on the
device), hidden activations tanh, output linear:

    a0 = bf(x)           z1 = a0 bf(W1)^T + b1    h1 = tanh(z1)
    a1 = bf(h1)          z2 = a1 bf(W2)^T + b2    h2 = tanh(z2)
    a2 = bf(h2)          z3 = a2 bf(W3)^T + b3    h3 = tanh(z3)
    a3 = bf(h3)          y  = a3 bf(W4)^T + b4
    loss = 1/2 mean_s |y_s - t_s|^2

Gradients treat the rounding as identity (straight-through), as the device does.  The
features follow the device's definition (differences of the history in float64 first).
There is no counterpart in the reference: the learned warm start is an addition of this
library; PCG/GMRES still run to the same tolerance."""
import numpy as np

NF, H, NO = 12, 64, 2
SHAPES = [("W1", (H, NF)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
          ("W4", (NO, H)), ("b4", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(x):
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta):
    theta = np.asarray(theta, np.float64)
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = theta[o:o + k].reshape(shp)
        o += k
    return out


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = []
    for name, shp in SHAPES:
        if name.startswith("W"):
            fan_in = shp[1]
            p.append(rng.normal(scale=1.0 / np.sqrt(fan_in), size=shp).ravel())
        else:
            p.append(np.zeros(shp))
    th = np.concatenate(p)
    # small output layer: the initial prediction is close to "no change"
    P = unpack(th)
    P["W4"] *= 0.1
    return th.astype(np.float32).astype(np.float64)


def identity(x):
    return np.asarray(x, np.float64)


def forward(theta, X, rnd=bf16):
    """rnd: the rounding of layer inputs and weights (bf16 as on the device; identity gives the
    exact network, used to pin the backward pass by finite differences)."""
    P = unpack(theta)
    a0 = rnd(X)
    h1 = np.tanh(a0 @ rnd(P["W1"]).T + P["b1"])
    a1 = rnd(h1)
    h2 = np.tanh(a1 @ rnd(P["W2"]).T + P["b2"])
    a2 = rnd(h2)
    h3 = np.tanh(a2 @ rnd(P["W3"]).T + P["b3"])
    a3 = rnd(h3)
    y = a3 @ rnd(P["W4"]).T + P["b4"]
    return y, (a0, h1, a1, h2, a2, h3, a3)


def loss_and_grad(theta, X, T, rnd=bf16):
    P = unpack(theta)
    S = X.shape[0]
    bf16 = rnd  # noqa: F811 -- the same rounding in the backward pass
    y, (a0, h1, a1, h2, a2, h3, a3) = forward(theta, X, rnd)
    d = y - T
    loss = 0.5 * np.sum(d * d) / S
    dy = d / S
    g = {}
    g["W4"] = dy.T @ a3
    g["b4"] = dy.sum(axis=0)
    dz3 = (dy @ bf16(P["W4"])) * (1.0 - h3 * h3)
    g["W3"] = dz3.T @ a2
    g["b3"] = dz3.sum(axis=0)
    dz2 = (dz3 @ bf16(P["W3"])) * (1.0 - h2 * h2)
    g["W2"] = dz2.T @ a1
    g["b2"] = dz2.sum(axis=0)
    dz1 = (dz2 @ bf16(P["W2"])) * (1.0 - h1 * h1)
    g["W1"] = dz1.T @ a0
    g["b1"] = dz1.sum(axis=0)
    return loss, np.concatenate([g[name].ravel() for name, _ in SHAPES])


def step_scale(h0, h1, n):
    """RMS over the DOFs of mu_n - mu_{n-1} and c_n - c_{n-1} (1 where it is 0)."""
    dm, dc = h0[:n] - h1[:n], h0[n:] - h1[n:]
    sm, sc = np.sqrt(np.sum(dm * dm) / n), np.sqrt(np.sum(dc * dc) / n)
    return (sm if sm > 0 else 1.0), (sc if sc > 0 else 1.0)


def _levels(hist, k):
    return list(hist) + [hist[-1]] * (k - len(hist))


def raw_features(hist, res, pcdiag, n):
    """hist: newest first, up to 4 states [mu; c] (2n each; missing levels repeat the oldest);
    res: residual b - K x_n of the new step's system, pcdiag its preconditioner diagonal (2n).
    Returns the 12 unnormalised features per DOF (float64) and the step scale."""
    h = _levels(hist, 4)
    sm, sc = step_scale(h[0], h[1], n)
    c = [v[n:] for v in h]
    mu = [v[:n] for v in h]
    F = np.empty((n, NF))
    F[:, 0] = c[0]
    F[:, 1] = (c[0] - c[1]) / sc
    F[:, 2] = (c[1] - c[2]) / sc
    F[:, 3] = (c[2] - c[3]) / sc
    F[:, 4] = mu[0]
    F[:, 5] = (mu[0] - mu[1]) / sm
    F[:, 6] = (mu[1] - mu[2]) / sm
    F[:, 7] = (mu[2] - mu[3]) / sm
    F[:, 8] = (res[:n] / pcdiag[:n]) / sm
    F[:, 9] = (res[n:] / pcdiag[n:]) / sc
    F[:, 10] = pcdiag[:n]
    F[:, 11] = pcdiag[n:]
    return F, (sm, sc)


def extrapolation(hist, n, scale):
    """dx_n/s per DOF and component (mu, c), float64: linear extrapolation in scaled units."""
    h = _levels(hist, 3)
    out = np.empty((n, NO))
    for o, (sl, s) in enumerate(((slice(0, n), scale[0]), (slice(n, 2 * n), scale[1]))):
        out[:, o] = (h[0][sl] - h[1][sl]) / s
    return out


def targets(x_next, hist, n, out_mean, out_scale):
    h = _levels(hist, 3)
    scale = step_scale(h[0], h[1], n)
    ex = extrapolation(h, n, scale)
    T = np.empty((n, NO))
    for o, (sl, s) in enumerate(((slice(0, n), scale[0]), (slice(n, 2 * n), scale[1]))):
        corr = (x_next[sl] - h[0][sl]) / s - ex[:, o]
        T[:, o] = (corr - out_mean[o]) / out_scale[o]
    return T
