"""GPU parity of the batched solve (many uneven P1 problems in one launch) through the C ABI.

Bars: per problem the CSR pattern, the operator-combination values and the load vector bit-exact
against the oracle (the batch runs the single-problem per-item code); the FP32-inner/FP64-refined
solution with true relative residual <= 1.5e-8 and within 1e-6 of the oracle's FP64 solution
solved to 1e-13; graph replays bit-identical to the eager launch; a problem with an exact-zero
node reports the reference's error without disturbing the others."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from drops_amd import batch as BT
from drops_amd import dhip as D
from drops_amd.mesh import KuhnMesh, sphere

pytestmark = pytest.mark.gpu

COEF = (1.0, 1.0, 1.0)


def lc_of(m):
    return (1.0, 1.0, m.h)


def small_set(nprob=24, seed=3):
    probs = BT.make_problems(nprob, seed, nmin=8, nmax=20)
    phi, off = BT.level_sets(probs)
    meshes = [p[0] for p in probs]
    return meshes, phi, off


def make_batch(ctx, meshes, phi, off, **kw):
    # each problem solves mass + LB + h * stabilisation with its own mesh size
    return ctx.batch(meshes, phi, off, coef=COEF, rhs_kind=D.RHS_3XY_MINUS_X3, **kw)


def test_batch_matches_oracle(ctx):
    print("[gpu batch] 24 problems, eager", flush=True)
    meshes, phi, off = small_set()
    b = make_batch(ctx, meshes, phi, off)
    b.solve(tol=1e-8)
    st = b.stats()
    assert np.all(st["status"] == D.OK), st["status"]
    assert np.all(st["resid"] <= 1e-8)
    assert len(set(st["ndof"].tolist())) > 5   # uneven sizes
    for p, m in enumerate(meshes):
        g = b.problem(p)
        sz = int(np.prod(m.halfgrid_shape))
        mm = KuhnMesh(m.orig, m.e, m.n)
        check_problem_lc(mm, phi[off[p]:off[p] + sz].reshape(m.halfgrid_shape), g, lc_of(m))
    b.close()


def check_problem_lc(m, phi, g, lc, tol=1e-8):
    om = O.make_mesh(m.orig, m.e, m.n)
    dof, nd = O.number_p1if(om, phi)
    r = O.assemble_p1(om, phi, dof, nd, coef=COEF, rhs_kind=O.RHS_3XY_MINUS_X3)
    comb = (1.0 * (lc[0] * r["mass"] + lc[1] * r["lb"])) + lc[2] * r["stab"]
    assert np.array_equal(g["rowptr"].astype(np.int64), r["rowptr"])
    assert np.array_equal(g["colind"], r["colind"])
    assert np.array_equal(g["comb"], comb)
    assert np.array_equal(g["rhs"], r["rhs"])
    A = sp.csr_matrix((comb, r["colind"], r["rowptr"]), shape=(nd, nd))
    b = r["rhs"]
    assert np.linalg.norm(b - A @ g["x"]) <= 1.5 * tol * np.linalg.norm(b)
    xo, it, res, ok = O.pcg_jacobi(r["rowptr"], r["colind"], comb, b, tol=1e-13, max_iter=100000)
    assert np.linalg.norm(g["x"] - xo) <= 1e-6 * np.linalg.norm(xo)


def test_batch_graph_replay_and_tiles(ctx):
    print("[gpu batch] graph replay, tile sizes", flush=True)
    meshes, phi, off = small_set(16, seed=5)
    b = make_batch(ctx, meshes, phi, off)
    b.solve(tol=1e-8)
    ref = [b.problem(p) for p in range(len(meshes))]
    ref_st = b.stats()
    for _ in range(3):
        b.solve(tol=1e-8, graph=True)
        st = b.stats()
        assert np.array_equal(st["iters"], ref_st["iters"])
        for p in range(len(meshes)):
            g = b.problem(p)
            assert np.array_equal(g["x"], ref[p]["x"])
    # other workgroup sizes: identical assembly, solutions to the same tolerance
    for threads in (128, 512):
        ctx.tune("batch_threads", threads)
        try:
            b.solve(tol=1e-8)
            st = b.stats()
            assert np.all(st["status"] == D.OK) and np.all(st["resid"] <= 1e-8)
            for p in range(len(meshes)):
                g = b.problem(p)
                assert np.array_equal(g["comb"], ref[p]["comb"])
                assert np.linalg.norm(g["x"] - ref[p]["x"]) <= 1e-6 * np.linalg.norm(ref[p]["x"])
        finally:
            ctx.tune("batch_threads", 256)
    # classification and local matrices by a grid-wide pre-pass instead of inside each problem's
    # workgroup: the same assembled systems, iteration counts and solutions bit for bit
    ctx.tune("batch_prepass", 1)
    try:
        for graph in (False, True):
            b.solve(tol=1e-8, graph=graph)
            st = b.stats()
            assert np.array_equal(st["iters"], ref_st["iters"]) and np.array_equal(st["ncut"], ref_st["ncut"])
            for p in range(len(meshes)):
                g = b.problem(p)
                for k in ("rowptr", "colind", "comb", "rhs", "x"):
                    assert np.array_equal(g[k], ref[p][k]), (p, k)
    finally:
        ctx.tune("batch_prepass", 0)
    # assembly launch and a separate PCG launch: with the same workgroup size bit for bit, with
    # 1024 threads per problem to the same tolerance
    ctx.tune("batch_split", 1)
    try:
        for sthreads in (256, 1024):
            ctx.tune("batch_solve_threads", sthreads)
            for graph in (False, True):
                b.solve(tol=1e-8, graph=graph)
                st = b.stats()
                assert np.all(st["status"] == D.OK) and np.all(st["resid"] <= 1e-8)
                for p in range(len(meshes)):
                    g = b.problem(p)
                    assert np.array_equal(g["comb"], ref[p]["comb"]) and np.array_equal(g["rhs"], ref[p]["rhs"])
                    if sthreads == 256:
                        assert np.array_equal(g["x"], ref[p]["x"]), p
                    else:
                        assert np.linalg.norm(g["x"] - ref[p]["x"]) <= 1e-6 * np.linalg.norm(ref[p]["x"])
    finally:
        ctx.tune("batch_split", 0)
        ctx.tune("batch_solve_threads", 1024)
    b.close()


def test_batch_zero_node_problem_isolated(ctx):
    print("[gpu batch] exact-zero node in one problem", flush=True)
    meshes, phi, off = small_set(4, seed=7)
    bad = KuhnMesh.box(-2.0, 2.0, 8)   # the unit sphere passes through grid vertices (1, 0, 0), ...
    pb = bad.sample_p2(sphere()).ravel()
    meshes = meshes + [bad]
    off = np.concatenate([off, [phi.size]])
    phi = np.concatenate([phi, pb])
    b = make_batch(ctx, meshes, phi, off)
    b.solve(tol=1e-8)
    st = b.stats()
    assert st["status"][-1] == D.E_ZERO_NODE
    assert np.all(st["status"][:-1] == D.OK)
    with pytest.raises(D.DhipError):
        b.problem(len(meshes) - 1)
    b.close()
