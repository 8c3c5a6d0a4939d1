"""GPU parity of the Krylov core (SpMV, dot, axpy, PCG) through the C ABI.

Bars (FP64): SpMV bit-identical to the oracle's serial row loop (the GPU sums each row
in column order, as the reference's y_Ax); dots within 1e-14 relative of the compensated
oracle dot; PCG solutions within 1e-8 relative of the oracle's solution solved to
1e-13, iteration counts within 2% of the oracle's PCG at the same tolerance.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import cases
import oracle as O
from drops_amd import dhip as D

pytestmark = pytest.mark.gpu


def assembled_system(ctx, n=32, lc_stab=None):
    mesh, phi = cases.c1_sphere(n)
    lc = (1.0, 1.0, mesh.h if lc_stab is None else lc_stab)
    print(f"[gpu krylov] system n={n}", flush=True)
    ctx.mesh(mesh)
    ctx.levelset(phi)
    ctx.assemble_p1(coef=(1.0, 1.0, 1.0), rhs_kind=D.RHS_3XY_MINUS_X3, lc=lc)
    g = ctx.csr()
    return g, ctx.csr_from_assembly(D.OP_COMB)


def test_spmv_dot_axpy(ctx):
    g, A = assembled_system(ctx, 24)
    rng = np.random.default_rng(0)
    n = g["rowptr"].size - 1
    x = rng.normal(size=n)
    y = ctx.spmv(A, x)
    rp = g["rowptr"].astype(np.int64)
    yr = O.spmv(rp, g["colind"], g["comb"], x)
    S = sp.csr_matrix((np.abs(g["comb"]), g["colind"], rp), shape=(n, n))
    assert np.all(np.abs(y - yr) <= 1e-15 * (S @ np.abs(x)) + 1e-300)
    # each row is summed in column order, as the oracle's (and the reference's y_Ax) serial loop
    assert np.array_equal(y, yr)
    d = ctx.dot(x, y)
    dr = O.kahan_dot(x, y)
    assert abs(d - dr) <= 1e-14 * (np.abs(x) @ np.abs(y))
    z = ctx.axpy(0.37, x, y)
    assert np.array_equal(z, y + 0.37 * x)
    A.close()


@pytest.mark.parametrize("variant", [D.FP64, D.FP64_CLASSIC])
@pytest.mark.parametrize("n", [16, 32, 64])
def test_pcg_fp64_matches_oracle(ctx, n, variant):
    """Both FP64 forms: the single-reduction arrangement (default) and the reference's
    two-reduction operation order."""
    g, A = assembled_system(ctx, n)
    rp = g["rowptr"].astype(np.int64)
    b = g["rhs"]
    x, it, res, ok = ctx.pcg(A, b, tol=1e-10, precision=variant)
    xr, itr, resr, okr = O.pcg_jacobi(rp, g["colind"], g["comb"], b, tol=1e-10)
    assert ok and okr and res <= 1e-10
    assert abs(it - itr) <= max(2, 0.02 * itr), (it, itr)
    xs, _, _, _ = O.pcg_jacobi(rp, g["colind"], g["comb"], b, tol=1e-13)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8
    # true residual of the GPU solution
    Am = sp.csr_matrix((g["comb"], g["colind"], rp), shape=(b.size, b.size))
    assert np.linalg.norm(b - Am @ x) / np.linalg.norm(b) < 2e-10
    A.close()


def test_pcg_mixed_precision(ctx):
    g, A = assembled_system(ctx, 64)
    rp = g["rowptr"].astype(np.int64)
    b = g["rhs"]
    x, it, res, ok = ctx.pcg(A, b, tol=1e-10, precision=D.FP32_IR)
    assert ok and res <= 1e-10
    xs, _, _, _ = O.pcg_jacobi(rp, g["colind"], g["comb"], b, tol=1e-13)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8
    A.close()


@pytest.mark.parametrize("variant", [D.FP64, D.FP64_CLASSIC, D.FP32_IR])
def test_pcg_not_converged_and_start_vector(ctx, variant):
    g, A = assembled_system(ctx, 16)
    b = g["rhs"]
    x, it, res, ok = ctx.pcg(A, b, tol=1e-10, max_iter=3, raise_on_fail=False, precision=variant)
    assert not ok and it == 3 and res > 1e-10
    # restart from the partial solution converges (x holds the start vector on entry)
    x2, it2, res2, ok2 = ctx.pcg(A, b, x0=x, tol=1e-10, precision=variant)
    assert ok2 and res2 <= 1e-10
    # zero right-hand side: converged at iteration 0
    x3, it3, res3, ok3 = ctx.pcg(A, np.zeros_like(b), tol=1e-10, rel=False, precision=variant)
    assert ok3 and it3 == 0 and not np.any(x3)
    A.close()


def test_pcg_on_uploaded_matrix(ctx):
    """A user-built CSR (MatrixCL via MatrixBuilderCL) with the unpreconditioned solver."""
    rng = np.random.default_rng(5)
    n = 300
    M = sp.random(n, n, density=0.02, random_state=5)
    M = (M + M.T) + sp.eye(n) * 4.0
    M = M.tocsr()
    M.sort_indices()
    b = rng.normal(size=n)
    A = ctx.csr_create(M.indptr, M.indices, M.data)
    for pc, variant in [(pc, v) for pc in (D.PC_NONE, D.PC_JACOBI) for v in (D.FP64, D.FP64_CLASSIC)]:
        x, it, res, ok = ctx.pcg(A, b, pc=pc, tol=1e-12, precision=variant)
        assert ok
        xd = np.linalg.solve(M.toarray(), b)
        assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-10
    A.close()


@pytest.mark.parametrize("variant", [D.FP64, D.FP64_CLASSIC, D.FP32_IR, "gmres"])
def test_xcd_row_mapping(ctx, variant):
    """Persistent solvers with each XCD's blocks owning one contiguous run of rows (xcd_map 1,
    the default) against the plain block order (xcd_map 0), on a block count that is not a
    multiple of 8 (47 rows per block): the same system solved to the same tolerance, the
    solutions within the solve's accuracy of each other and of the oracle's tight solution."""
    g, A = assembled_system(ctx, 48)
    rp = g["rowptr"].astype(np.int64)
    b = g["rhs"]
    n = b.size
    assert (n + 46) // 47 % 8 != 0
    xs, _, _, _ = O.pcg_jacobi(rp, g["colind"], g["comb"], b, tol=1e-13)
    out = {}
    try:
        ctx.tune("pcg_rows_per_block", 47)
        for xm in (0, 1):
            ctx.tune("xcd_map", xm)
            print(f"[gpu krylov] xcd_map={xm} variant={variant}", flush=True)
            if variant == "gmres":
                x, it, res, ok = ctx.gmres(A, b, tol=1e-10)
            else:
                x, it, res, ok = ctx.pcg(A, b, tol=1e-10, precision=variant)
            assert ok and res <= 1e-10
            assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8
            out[xm] = (x, it)
    finally:
        ctx.tune("pcg_rows_per_block", 0)
        ctx.tune("xcd_map", 1)
    (x0, it0), (x1, it1) = out[0], out[1]
    assert abs(it0 - it1) <= max(2, 0.02 * it0)
    assert np.linalg.norm(x0 - x1) / np.linalg.norm(x0) < 1e-8
    A.close()


def test_profile_solver_only(ctx):
    """profile mode 2 brackets only the persistent solver launches with events; mode 1 every launch"""
    mesh, phi = cases.c1_sphere(24)
    ctx.mesh(mesh)
    ctx.levelset(phi)
    try:
        for mode, expect_classify in ((2, 0), (1, 1)):
            ctx.profile_reset()
            ctx.profile(mode)
            ctx.assemble_p1(coef=(1.0, 1.0, 1.0), rhs_kind=D.RHS_3XY_MINUS_X3, lc=(1.0, 1.0, mesh.h))
            A = ctx.csr_from_assembly(D.OP_COMB)
            ctx.pcg(A, ctx.csr()["rhs"], tol=1e-10)
            A.close()
            ms, cnt = ctx.profile_get("pcg64")
            assert cnt == 1 and ms > 0.0
            assert ctx.profile_get("classify")[1] == expect_classify
        with pytest.raises(D.DhipError) as e:
            ctx.profile(3)
        assert e.value.code == D.E_ARG
    finally:
        ctx.profile(0)


def test_spmv_row_lengths_and_fallback(ctx):
    """SpMV tiles whose entries fit the LDS stage and tiles that do not (rows of 0 to 400
    entries, the long ones grouped so that whole tiles overflow): every row summed in column
    order, bit-identical to the oracle's serial loop."""
    rng = np.random.default_rng(5)
    n = 3000
    lens = rng.integers(0, 20, size=n)
    lens[1000:1100] = rng.integers(200, 400, size=100)   # tiles of these rows exceed the stage
    lens[2000:2010] = 0
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(n, size=k, replace=False)) for k in lens]).astype(np.int32)
    val = rng.normal(size=col.size)
    x = rng.normal(size=n)
    print(f"[gpu krylov] spmv ragged n={n} nnz={col.size}", flush=True)
    A = ctx.csr_create(rp.astype(np.int32), col, val)
    y = ctx.spmv(A, x)
    A.close()
    assert np.array_equal(y, O.spmv(rp, col, val, x))
    assert np.all(y[2000:2010] == 0.0)
