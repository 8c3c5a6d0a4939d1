"""File formats of device data: an assembled operator exported from its device handle
(MatrixCL::exportMTX / out) is byte-identical to the iostream restatement of the same
statements on the oracle's matrix; a matrix file imported into a device handle gives the
same SpMV; device vectors written and read back in both vector formats."""
import numpy as np
import pytest

import cases
import oracle as O
from drops_amd import dhip as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt", ["mtx", "plain"])
def test_export_assembled_operator(ctx, tmp_path, fmt):
    mesh, phi = cases.PARITY_CASES["c1_sphere16"]()
    print(f"[gpu formats] export {fmt} n={mesh.n}", flush=True)
    ctx.mesh(mesh)
    ctx.levelset(phi)
    lc = (1.0, 1.0, 0.25)
    ctx.assemble_p1(coef=(1.0, 1.0, 1.0), rhs_kind=D.RHS_3XY_MINUS_X3, lc=lc)
    A = ctx.csr_from_assembly(D.OP_COMB)
    ours = str(tmp_path / "ours")
    A.export(ours, fmt)
    A.close()
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    dof, nd = O.number_p1if(om, phi)
    ref = O.assemble_p1(om, phi, dof, nd, coef=(1.0, 1.0, 1.0), rhs_kind=O.RHS_3XY_MINUS_X3)
    comb = (1.0 * (lc[0] * ref["mass"] + lc[1] * ref["lb"])) + lc[2] * ref["stab"]
    path = str(tmp_path / "ref")
    O.fmt_matrix(path, ref["rowptr"], ref["colind"], comb, fmt=fmt)
    with open(ours, "rb") as a, open(path, "rb") as b:
        assert a.read() == b.read()


@pytest.mark.parametrize("fmt", ["mtx", "plain"])
def test_import_spmv(ctx, tmp_path, fmt):
    rng = np.random.default_rng(11)
    n = 5000
    rows = [np.sort(rng.choice(n, size=int(rng.integers(0, 12)), replace=False)) for _ in range(n)]
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate(rows)
    # values with 6 significant digits survive the plain format exactly as written
    v = np.array([float("%g" % x) for x in rng.standard_normal(ci.size)])
    path = str(tmp_path / "m")
    O.fmt_matrix(path, rp, ci, v, fmt=fmt)
    print(f"[gpu formats] import {fmt} n={n} nnz={ci.size}", flush=True)
    A = ctx.csr_import(path, fmt)
    assert A.shape == (n, ci.size)
    # back out of the device handle: the same bytes
    again = str(tmp_path / "again")
    A.export(again, fmt)
    with open(path, "rb") as a, open(again, "rb") as b:
        assert a.read() == b.read()
    x = rng.standard_normal(n)
    y = ctx.spmv(A, x)
    A.close()
    y_ref = O.spmv(rp, ci.astype(np.int32), v, x)
    # the device SpMV sums a row in a lane-split order: within the rounding bound of the row
    mag = O.spmv(rp, ci.astype(np.int32), np.abs(v), np.abs(x))
    assert np.all(np.abs(y - y_ref) <= 2e-15 * mag + 1e-300)


def test_device_vector_round_trip(ctx, tmp_path):
    rng = np.random.default_rng(12)
    x = rng.standard_normal(100000) * 10.0 ** rng.integers(-200, 200, 100000)
    p = ctx.to_device(x)
    q = ctx.malloc(x.nbytes)
    try:
        for binary in (True, False):
            print(f"[gpu formats] vector binary={binary}", flush=True)
            path = str(tmp_path / f"v{int(binary)}")
            ctx.vec_write(path, p, x.size, binary=binary)
            ref = str(tmp_path / f"r{int(binary)}")
            O.fmt_vector(ref, x, binary=binary)
            with open(path, "rb") as a, open(ref, "rb") as b:
                assert a.read() == b.read()
            ctx.vec_read(path, q, x.size, binary=binary)
            back = ctx.to_host(q, x.shape, np.float64)
            if binary:
                assert np.array_equal(back.view(np.uint64), x.view(np.uint64))
            else:
                assert np.array_equal(back, O.read_vector(ref, x.size, binary=False))
        with pytest.raises(D.DhipError) as e:
            ctx.vec_read(str(tmp_path / "v1"), q, x.size - 1, binary=True)
        assert e.value.code == D.E_ARG
    finally:
        ctx.free(p)
        ctx.free(q)
