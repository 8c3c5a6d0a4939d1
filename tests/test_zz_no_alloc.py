"""The allocation-free PCG (dhip_pcg_noalloc) against dhip_pcg with precision DHIP_FP64.

Both run the same persistent kernel at the same launch geometry; the allocation-free entry point
keeps its diagonal, work vectors, slots and result record in the library's static scratch array.
The solutions, iteration counts and residuals must therefore be bit-identical, on two calls in a
row (the second on a smaller system, so that stale scratch contents of the first call would show),
and a system above DHIP_NOALLOC_MAX_ROWS rows must be rejected before any launch."""
import ctypes as C

import numpy as np
import pytest

from drops_amd import dhip as D
from drops_amd.mesh import KuhnMesh, config_shift, sphere

pytestmark = pytest.mark.gpu

NOALLOC_MAX_ROWS = 1048576


def _system(ctx, n):
    mesh = KuhnMesh.box(-2.0, 2.0, n, shift=config_shift(4.0 / n))
    ctx.mesh(mesh)
    ctx.levelset(mesh.sample_p2(sphere()))
    ctx.assemble_p1(coef=(1.0, 1.0, 1.0), rhs_kind=D.RHS_3XY_MINUS_X3, lc=(1.0, 1.0, mesh.h))
    g = ctx.csr()
    return ctx.csr_from_assembly(D.OP_COMB), g["rhs"]


def _solve_noalloc(ctx, A, b, tol):
    bd = ctx.to_device(b)
    xd = ctx.to_device(np.zeros(b.size))
    it = C.c_int32(5000)
    t = C.c_double(tol)
    print(f"dhip_pcg_noalloc rows={b.size} tol={tol}", flush=True)
    rc = ctx.L.dhip_pcg_noalloc(ctx.h, A.h, D.PC_JACOBI, C.c_void_p(bd), C.c_void_p(xd), C.byref(it), C.byref(t), 1)
    x = ctx.to_host(xd, b.size, np.float64)
    ctx.free(bd)
    ctx.free(xd)
    return rc, x, it.value, t.value


@pytest.mark.parametrize("sizes", [(32, 16)])
def test_noalloc_matches_pcg_twice(sizes):
    with D.Context(0) as ctx:
        for n in sizes:
            A, b = _system(ctx, n)
            x_ref, it_ref, res_ref, ok = ctx.pcg(A, b, tol=1e-10, precision=D.FP64)
            assert ok
            rc, x, it, res = _solve_noalloc(ctx, A, b, 1e-10)
            assert rc == D.OK, ctx.L.dhip_last_error(ctx.h)
            assert it == it_ref and res == res_ref, (n, it, it_ref, res, res_ref)
            assert np.array_equal(x, x_ref), (n, np.abs(x - x_ref).max())
            A.close()


def test_noalloc_rejects_oversized_system():
    with D.Context(0) as ctx:
        n = NOALLOC_MAX_ROWS + 1
        rp = np.arange(n + 1, dtype=np.int32)
        col = np.arange(n, dtype=np.int32)
        A = ctx.csr_create(rp, col, np.ones(n))
        b = np.ones(n)
        rc, _, _, _ = _solve_noalloc(ctx, A, b, 1e-10)
        assert rc == D.E_CAPACITY
        assert b"max 1048576 rows" in ctx.L.dhip_last_error(ctx.h)
        A.close()
