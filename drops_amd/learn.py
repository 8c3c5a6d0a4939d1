"""Learned warm start for the Cahn-Hilliard trajectories: data from the device stepper, training
with the library's own forward/backward kernels, evaluation against the previous-step start.

One sample per DOF and step n -> n+1: the 12 features of dhip_mlp_features (last four time levels
of [mu; c], the step's right-hand side and preconditioner diagonal) and the normalised change
of [mu; c].  Everything stays on the device; only normalisation statistics (from a subsample)
and scalar losses come back to the host."""
import json
import os
import time

import numpy as np

from . import ch_run
from . import dhip as D
from .mlp_init import init_params

NF, NO = D.MLP_NF, 2


class Recorder:
    """Device copies of every state, right-hand side and preconditioner diagonal of a trajectory."""

    def __init__(self, ctx, ch, steps):
        self.ctx, self.ch, self.steps, self.n = ctx, ch, steps, ch.n
        n2 = 2 * ch.n
        self.bytes = 8 * n2
        self.states = ctx.malloc(self.bytes * (steps + 1))
        self.res = ctx.malloc(self.bytes * steps)   # b - K x_n of each step's system
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
            ch.residual(self.res + s * self.bytes)
            ctx.copy_d2d(self.pcd + s * self.bytes, sysp["pcdiag"], self.bytes)
            its[s], _, _ = ch.solve()
            ctx.copy_d2d(self.state(s + 1), sysp["state"], self.bytes)
        return its

    def hist(self, s, levels=4):
        """history pointers for step s -> s+1, newest first (levels before the start repeat it)"""
        return [self.state(s - k) for k in range(min(levels, s + 1))]

    def free(self):
        for p in (self.states, self.res, self.pcd):
            self.ctx.free(p)


def normalisation(ctx, mlp, recs, every=5, max_samples=4_000_000):
    """Feature mean/scale and target mean/scale from a subsample of steps (and of DOFs, so that at
    most about max_samples rows come to the host)."""
    n = recs[0].n
    mlp.set_normalization(np.zeros(NF), np.ones(NF), np.zeros(NO), np.ones(NO))
    Xd = ctx.malloc(4 * NF * n)
    Td = ctx.malloc(4 * NO * n)
    feats, targs = [], []
    nsel = len(recs) * len(range(1, recs[0].steps, every))
    stride = max(1, (nsel * n + max_samples - 1) // max_samples)
    for r in recs:
        for s in range(1, r.steps, every):
            mlp.features(r.hist(s), r.res + s * r.bytes, r.pcd + s * r.bytes, n, Xd)
            mlp.targets(r.state(s + 1), r.hist(s, 3), n, Td)
            feats.append(ctx.to_host(Xd, (n, NF), np.float32)[::stride].astype(np.float64))
            targs.append(ctx.to_host(Td, (n, NO), np.float32)[::stride].astype(np.float64))
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
        per = n * (recs[0].steps - 1)
        self.S = per * len(recs)
        self.X = ctx.malloc(4 * NF * self.S)
        self.T = ctx.malloc(4 * NO * self.S)
        o = 0
        for r in recs:
            for s in range(1, r.steps):   # step 0 has no history and always starts from x_0
                mlp.features(r.hist(s), r.res + s * r.bytes, r.pcd + s * r.bytes, n, self.X + 4 * NF * o)
                mlp.targets(r.state(s + 1), r.hist(s, 3), n, self.T + 4 * NO * o)
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


def null_network(ctx, mlp):
    """A network whose correction is zero: the learned path then starts from x_n or the linear
    extrapolation only (the method without the trained network)."""
    null = ctx.mlp(np.zeros(D.MLP_NPARAM, np.float32))
    null.set_normalization(*[np.asarray(v) for v in mlp.norm[:2]], np.zeros(NO), np.ones(NO))
    return null


def evaluate(ch, mlp, seed, steps, variants=("previous", "extrapolation", "learned"), first=1):
    """The same trajectory from the previous-step start, from the linear extrapolation alone
    (zero network) and from the learned start; a "_mixed" suffix runs the variant with FP32
    Arnoldi cycles inside FP64 restarts (dhip_ch_set_precision FP32_IR).  Per-step iterations and
    wall time (prepare + solve) of each, the final state and the state after `first` steps."""
    out = {}
    null = null_network(ch.ctx, mlp) if "extrapolation" in variants and mlp is not None else None
    for name in variants:
        base = name[:-len("_mixed")] if name.endswith("_mixed") else name
        net = null if base == "extrapolation" else mlp
        ch.set_precision(D.FP32_IR if name.endswith("_mixed") else D.FP64)
        c0, mu0 = ch_run.initial_state(ch.n, seed)
        ch.set_state(c0, mu0)
        its = np.zeros(steps, np.int32)
        wall = np.zeros(steps)
        used = [0, 0, 0]
        state_first = None
        for s in range(steps):
            t0 = time.perf_counter()
            ch.prepare()
            if base != "previous":
                its[s], _, _, u = ch.solve_learned(net)
                used[u] += 1
            else:
                its[s], _, _ = ch.solve()
            wall[s] = time.perf_counter() - t0
            if s + 1 == first:
                state_first = ch.state()
        out[name] = dict(iters=its, wall=wall, state=ch.state(), state_first=state_first,
                         starts=dict(previous=used[0], network=used[1], extrapolation=used[2]))
    ch.set_precision(D.FP64)
    if null is not None:
        null.close()
    return out


DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def save(path, mlp, norm, meta):
    np.savez(path, theta=mlp.params(), in_mean=norm[0], in_scale=norm[1], out_mean=norm[2], out_scale=norm[3],
             meta=json.dumps(meta))


def load(ctx, path):
    """A trained network from save(); returns (mlp, meta)."""
    z = np.load(path, allow_pickle=False)
    mlp = ctx.mlp(z["theta"])
    mlp.set_normalization(z["in_mean"], z["in_scale"], z["out_mean"], z["out_scale"])
    return mlp, json.loads(str(z["meta"]))


def default_weights(space):
    return os.path.join(DATA_DIR, "ch_warmstart_torus96_p2.npz" if space == 2 else "ch_warmstart_torus96.npz")


def main(argv=None):
    """Train the warm-start network on seeded torus trajectories and report its effect on unseen
    seeds:  python -m drops_amd.learn [--space 2] [--cells 96] [--steps 200] [--train-seeds 20]
    [--test-seeds 5] [--iters 3000] [--out drops_amd/data/ch_warmstart_torus96_p2.npz]"""
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, default=96)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--train-seeds", type=int, default=20)
    ap.add_argument("--test-seeds", type=int, default=5)
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--space", type=int, default=2, choices=(1, 2))
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    if a.out is None:
        a.out = default_weights(a.space)
    log = lambda *x: print(*x, flush=True)
    m, phi = ch_run.torus_mesh(a.cells)
    with D.Context(0) as ctx:
        ch = ch_run.make_stepper(ctx, m, phi, space=a.space)
        log(f"[learn] torus {a.cells}^3, P{a.space}: {ch.n} DOFs per field, {a.steps} steps")
        t0 = time.perf_counter()
        recs = []
        for seed in range(a.train_seeds):
            r = Recorder(ctx, ch, a.steps)
            its = r.run(seed)
            recs.append(r)
            log(f"[learn] trajectory seed {seed}: {int(its.sum())} GMRES iterations")
        t_data = time.perf_counter() - t0
        mlp = ctx.mlp(init_params(0))
        norm = normalisation(ctx, mlp, recs)
        data = Dataset(ctx, mlp, recs)
        for r in recs:
            r.free()
        t0 = time.perf_counter()
        hist = train(ctx, mlp, data, iters=a.iters, batch=a.batch, log_every=max(1, a.iters // 10))
        ctx.synchronize()
        t_train = time.perf_counter() - t0
        for it, loss in hist:
            log(f"[learn] iter {it}: loss {loss:.4e}")
        data.free()
        res = []
        for seed in range(1000, 1000 + a.test_seeds):
            ev = evaluate(ch, mlp, seed, a.steps)
            p, e, q = ev["previous"], ev["extrapolation"], ev["learned"]
            err = float(np.linalg.norm(p["state"][0] - q["state"][0]) / np.linalg.norm(p["state"][0]))
            res.append(dict(seed=seed, iters_previous=int(p["iters"].sum()), iters_extrapolation=int(e["iters"].sum()),
                            iters_learned=int(q["iters"].sum()), s_previous=float(p["wall"].sum()),
                            s_extrapolation=float(e["wall"].sum()), s_learned=float(q["wall"].sum()),
                            starts_learned=q["starts"], c_rel_diff=err))
            log(f"[learn] test seed {seed}: {res[-1]}")
        meta = dict(space=a.space, cells=a.cells, steps=a.steps, train_seeds=list(range(a.train_seeds)), dt=ch_run.DT,
                    epsilon=ch_run.EPS, iters=a.iters, batch=a.batch, final_loss=hist[-1][1], data_s=t_data,
                    train_s=t_train, test=res)
        os.makedirs(os.path.dirname(a.out), exist_ok=True)
        save(a.out, mlp, norm, meta)
        log(json.dumps(dict(out=a.out, train_s=t_train, data_s=t_data, final_loss=hist[-1][1], test=res)))
        mlp.close()
        ch.close()


if __name__ == "__main__":
    main()
