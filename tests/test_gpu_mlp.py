"""GPU checks of the learned warm start against the CPU oracle (oracle/mlp_oracle.py):
features bit-identical, minibatch gradients and losses of the device training kernel within
1e-3 (float32 accumulation against float64, identical bfloat16 rounding points), matrix-core
inference within 2e-3 of the oracle's forward pass, training that lowers the loss, and a
trained network that cuts GMRES iterations against the previous-step start on an unseen seed."""
import numpy as np
import pytest

import mlp_oracle as M
from drops_amd import ch_run, learn
from drops_amd import dhip as D

pytestmark = pytest.mark.gpu


def upload(ctx, a):
    return ctx.to_device(np.ascontiguousarray(a))


def test_features_match_oracle(ctx):
    """Features within float32 rounding of the oracle's (the device's step scale sums its squares
    in block order, so the scale may differ from the oracle's in the last bits); targets likewise."""
    print("[gpu mlp] features", flush=True)
    rng = np.random.default_rng(0)
    n = 777
    hist = [rng.normal(size=2 * n) for _ in range(3)]
    xn = rng.normal(size=2 * n)
    res = rng.normal(size=2 * n)
    pcd = rng.uniform(0.5, 2.0, 2 * n)
    mean, scale = rng.normal(size=12), rng.uniform(0.5, 3.0, 12)
    omean, oscale = rng.normal(size=2), rng.uniform(0.5, 2.0, 2)
    mlp = ctx.mlp()
    mlp.set_normalization(mean, scale, omean, oscale)
    hp = [upload(ctx, h) for h in hist]
    rp, pp, xp = upload(ctx, res), upload(ctx, pcd), upload(ctx, xn)
    Xd = ctx.malloc(4 * 12 * n)
    Td = ctx.malloc(4 * 2 * n)
    for nh in (1, 2, 3):
        mlp.features(hp[:nh], rp, pp, n, Xd)
        X = ctx.to_host(Xd, (n, 12), np.float32)
        F, _ = M.raw_features(hist[:nh], res, pcd, n)
        ref = ((F - mean) * scale).astype(np.float32)
        assert np.all(np.abs(X - ref) <= 2.0 ** -21 * np.maximum(np.abs(ref), 1.0)), nh
        mlp.targets(xp, hp[:nh], n, Td)
        T = ctx.to_host(Td, (n, 2), np.float32)
        Tr = M.targets(xn, hist[:nh], n, omean, oscale).astype(np.float32)
        assert np.all(np.abs(T - Tr) <= 1e-5 * np.maximum(np.abs(Tr), 1.0)), nh
    for p in hp + [rp, pp, xp, Xd, Td]:
        ctx.free(p)
    mlp.close()


def test_grad_and_inference_match_oracle(ctx):
    print("[gpu mlp] gradient / inference", flush=True)
    rng = np.random.default_rng(1)
    th = M.init_params(5) + rng.normal(scale=0.05, size=M.NPARAM)
    th = th.astype(np.float32).astype(np.float64)
    S = 1000
    X = rng.normal(size=(S, 12)).astype(np.float32)
    T = (0.5 * rng.normal(size=(S, 2))).astype(np.float32)
    mlp = ctx.mlp(th)
    Xd, Td = upload(ctx, X), upload(ctx, T)
    idx = np.arange(S, dtype=np.int32)
    idxd = upload(ctx, idx)
    loss, g = mlp.grad(Xd, Td, idxd, S)
    lo, go = M.loss_and_grad(th, X.astype(np.float64), T.astype(np.float64))
    assert abs(loss - lo) <= 1e-4 * lo
    assert np.linalg.norm(g - go) <= 1e-3 * np.linalg.norm(go)
    # a permuted minibatch of half the samples
    sub = rng.permutation(S)[:500].astype(np.int32)
    subd = upload(ctx, sub)
    loss2, g2 = mlp.grad(Xd, Td, subd, 500)
    lo2, go2 = M.loss_and_grad(th, X[sub].astype(np.float64), T[sub].astype(np.float64))
    assert np.linalg.norm(g2 - go2) <= 1e-3 * np.linalg.norm(go2)
    # matrix-core inference
    Yd = ctx.malloc(4 * 2 * S)
    mlp.predict(Xd, S, Yd)
    Y = ctx.to_host(Yd, (S, 2), np.float32)
    yo, _ = M.forward(th, X.astype(np.float64))
    assert np.linalg.norm(Y - yo) <= 2e-3 * np.linalg.norm(yo)
    for p in (Xd, Td, idxd, subd, Yd):
        ctx.free(p)
    mlp.close()


def test_training_lowers_loss(ctx):
    print("[gpu mlp] training", flush=True)
    rng = np.random.default_rng(2)
    teacher = M.init_params(9) * 2.0
    S = 20000
    X = rng.normal(size=(S, 12))
    T, _ = M.forward(teacher, X, rnd=M.identity)
    T = (T / T.std(axis=0)).astype(np.float32)
    mlp = ctx.mlp(M.init_params(3))
    Xd, Td = upload(ctx, X.astype(np.float32)), upload(ctx, T)

    class Data:
        pass
    data = Data()
    data.X, data.T, data.S = Xd, Td, S
    hist = learn.train(ctx, mlp, data, iters=400, batch=4096, lr=3e-3, lr_end=3e-4, log_every=50)
    assert hist[-1][1] < 0.25 * hist[0][1], hist
    assert all(b[1] < a[1] * 1.05 for a, b in zip(hist, hist[1:])), hist
    # the device parameters are the ones the oracle evaluates to the same loss
    th = mlp.params().astype(np.float64)
    lo, _ = M.loss_and_grad(th, X.astype(np.float32).astype(np.float64), T.astype(np.float64))
    loss, _ = mlp.grad(Xd, Td, upload(ctx, np.arange(S, dtype=np.int32)), S)
    assert abs(loss - lo) <= 1e-3 * lo
    ctx.free(Xd)
    ctx.free(Td)
    mlp.close()


def test_learned_start_cuts_iterations(ctx):
    """Train on four short trajectories of a 24^3 torus, evaluate on an unseen seed."""
    print("[gpu mlp] learned start on torus n=24", flush=True)
    m, phi = ch_run.torus_mesh(24)
    ch = ch_run.make_stepper(ctx, m, phi)
    steps = 40
    recs = []
    for seed in range(4):
        r = learn.Recorder(ctx, ch, steps)
        r.run(seed)
        recs.append(r)
    mlp = ctx.mlp(M.init_params(0))
    learn.normalisation(ctx, mlp, recs)
    data = learn.Dataset(ctx, mlp, recs)
    learn.train(ctx, mlp, data, iters=600, batch=8192, lr=2e-3, lr_end=1e-4)
    res = learn.evaluate(ch, mlp, seed=100, steps=steps)
    prev, lrn = res["previous"]["iters"], res["learned"]["iters"]
    print(f"[gpu mlp] iterations previous-step {prev.sum()} learned {lrn.sum()} "
          f"(starts {res['learned']['starts']})", flush=True)
    assert lrn.sum() < 0.7 * prev.sum()
    assert res["learned"]["starts"]["network"] > 0
    # the same discrete trajectory either way: every step solved to 1e-8, 40 steps accumulate
    # differences at the level of the tolerance times the conditioning
    cp, cl = res["previous"]["state"][0], res["learned"]["state"][0]
    assert np.linalg.norm(cp - cl) <= 1e-5 * np.linalg.norm(cp)
    data.free()
    for r in recs:
        r.free()
    mlp.close()
    ch.close()
