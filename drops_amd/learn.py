"""Learned warm start for the Cahn-Hilliard trajectories: data from the device stepper, training
with the library's own forward/backward kernels, evaluation against the previous-step start.

One sample per DOF and step n -> n+1: the 12 features of dhip_mlp_features (last four time levels
of [mu; c], the step's right-hand side and preconditioner diagonal) and the normalised change
of [mu; c].  Everything stays on the device; only normalisation statistics (from a subsample)
and scalar losses come back to the host."""
import time

import numpy as np

from . import ch_run
from . import dhip as D

NF, NO = D.MLP_NF, 2


class Recorder:
    """Device copies of every state, right-hand side and preconditioner diagonal of a trajectory."""

    def __init__(self, ctx, ch, steps):
        self.ctx, self.ch, self.steps, self.n = ctx, ch, steps, ch.n
        n2 = 2 * ch.n
        self.bytes = 8 * n2
        self.states = ctx.malloc(self.bytes * (steps + 1))
        self.rhs = ctx.malloc(self.bytes * steps)
        self.pcd = ctx.malloc(self.bytes * steps)

    def state(self, k):
        return self.states + k * self.bytes

    def run(self, seed):
        ch, ctx = self.ch, self.ctx
        c0, mu0 = ch_run.initial_state(ch.n, seed)
        ch.set_state(c0, mu0)
        sysp = ch.system()
        ctx.copy_d2d(self.state(0), sysp["state"], self.bytes)
        its = np.zeros(self.steps, np.int32)
        for s in range(self.steps):
            ch.prepare()
            ctx.copy_d2d(self.rhs + s * self.bytes, sysp["rhs"], self.bytes)
            ctx.copy_d2d(self.pcd + s * self.bytes, sysp["pcdiag"], self.bytes)
            its[s], _, _ = ch.solve()
            ctx.copy_d2d(self.state(s + 1), sysp["state"], self.bytes)
        return its

    def hist(self, s):
        """history pointers for step s -> s+1, newest first"""
        return [self.state(max(s - k, 0)) for k in range(4)]

    def free(self):
        for p in (self.states, self.rhs, self.pcd):
            self.ctx.free(p)


def normalisation(ctx, mlp, recs, every=5):
    """Feature mean/scale and target mean/scale from a subsample of steps."""
    n = recs[0].n
    mlp.set_normalization(np.zeros(NF), np.ones(NF), np.zeros(NO), np.ones(NO))
    Xd = ctx.malloc(4 * NF * n)
    Td = ctx.malloc(4 * NO * n)
    feats, targs = [], []
    for r in recs:
        for s in range(0, r.steps, every):
            mlp.features(r.hist(s), r.rhs + s * r.bytes, r.pcd + s * r.bytes, n, Xd)
            mlp.targets(r.state(s + 1), r.state(s), n, Td)
            feats.append(ctx.to_host(Xd, (n, NF), np.float32).astype(np.float64))
            targs.append(ctx.to_host(Td, (n, NO), np.float32).astype(np.float64))
    ctx.free(Xd)
    ctx.free(Td)
    F, T = np.concatenate(feats), np.concatenate(targs)
    in_mean, in_std = F.mean(axis=0), F.std(axis=0)
    in_std[in_std == 0] = 1.0
    out_mean, out_std = T.mean(axis=0), T.std(axis=0)
    out_std[out_std == 0] = 1.0
    norm = (in_mean, 1.0 / in_std, out_mean, out_std)
    mlp.set_normalization(*norm)
    return norm


class Dataset:
    """Normalised features X [S][12] and targets T [S][2] (float32, device) of all recorded steps."""

    def __init__(self, ctx, mlp, recs):
        self.ctx = ctx
        n = recs[0].n
        per = n * recs[0].steps
        self.S = per * len(recs)
        self.X = ctx.malloc(4 * NF * self.S)
        self.T = ctx.malloc(4 * NO * self.S)
        o = 0
        for r in recs:
            for s in range(r.steps):
                mlp.features(r.hist(s), r.rhs + s * r.bytes, r.pcd + s * r.bytes, n, self.X + 4 * NF * o)
                mlp.targets(r.state(s + 1), r.state(s), n, self.T + 4 * NO * o)
                o += n

    def free(self):
        self.ctx.free(self.X)
        self.ctx.free(self.T)


def train(ctx, mlp, data, iters=2000, batch=65536, lr=2e-3, lr_end=1e-4, seed=0, log_every=0):
    """Adam on random minibatches; cosine decay lr -> lr_end.  Returns the loss history."""
    rng = np.random.default_rng(seed)
    idx_d = ctx.malloc(4 * batch)
    losses = []
    for it in range(iters):
        idx = rng.integers(0, data.S, size=batch, dtype=np.int64).astype(np.int32)
        ctx.check(ctx.L.dhip_memcpy(ctx.h, D.C.c_void_p(idx_d), D._vp(idx), idx.nbytes, 0))
        a = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * it / max(iters - 1, 1)))
        want = log_every and (it % log_every == 0 or it == iters - 1)
        loss = mlp.train_step(data.X, data.T, idx_d, batch, float(a), want_loss=bool(want))
        if want:
            losses.append((it, loss))
    ctx.free(idx_d)
    return losses


def evaluate(ch, mlp, seed, steps):
    """Same trajectory twice: previous-step start and learned start.  Per-step iterations and
    wall time (prepare + solve) of each."""
    out = {}
    for name in ("previous", "learned"):
        c0, mu0 = ch_run.initial_state(ch.n, seed)
        ch.set_state(c0, mu0)
        its = np.zeros(steps, np.int32)
        wall = np.zeros(steps)
        for s in range(steps):
            t0 = time.perf_counter()
            ch.prepare()
            if name == "learned":
                its[s], _, _ = ch.solve_learned(mlp)
            else:
                its[s], _, _ = ch.solve()
            wall[s] = time.perf_counter() - t0
        out[name] = dict(iters=its, wall=wall, state=ch.state())
    return out
