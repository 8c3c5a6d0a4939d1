"""GPU tests of y = A^T x (dhip_spmv_t, transp_mul / y_ATx of num/spmat.h:1729-1757) through the C
ABI: bit for bit against y_ATx's own loop (y_j += a_ij x_i over rows i in order) on rectangular and
square matrices with empty rows and columns, and a column index beyond ncols refused."""
import numpy as np
import pytest

from drops_amd import dhip as D

pytestmark = pytest.mark.gpu


def y_atx(rp, col, val, x, ncols):
    """num/spmat.h:1729-1747 restated: y_j += a_ij * x_i, rows in order"""
    y = np.zeros(ncols)
    for i in range(rp.size - 1):
        for q in range(rp[i], rp[i + 1]):
            y[col[q]] += val[q] * x[i]
    return y


def random_csr(rng, n, m, per_row):
    rp = [0]
    cols = []
    for i in range(n):
        k = 0 if i % 7 == 3 else rng.integers(1, per_row + 1)
        c = np.sort(rng.choice(m - 5, size=min(k, m - 5), replace=False))   # the last 5 columns stay empty
        cols.append(c)
        rp.append(rp[-1] + c.size)
    col = np.concatenate(cols).astype(np.int32)
    return np.array(rp, np.int32), col, rng.normal(size=col.size)


@pytest.mark.parametrize("n,m,per_row", [(300, 200, 12), (257, 257, 40), (64, 1000, 3)])
def test_spmv_t_bit_exact(ctx, n, m, per_row):
    print(f"[gpu spmv_t] n={n} m={m}", flush=True)
    rng = np.random.default_rng(n + m)
    rp, col, val = random_csr(rng, n, m, per_row)
    A = ctx.csr_create(rp, col, val)
    x = rng.normal(size=n)
    y = ctx.spmv_t(A, x, m)
    assert np.array_equal(y, y_atx(rp, col, val, x, m))
    assert np.all(y[m - 5:] == 0.0)
    with pytest.raises(D.DhipError) as ei:
        ctx.spmv_t(A, x, int(col.max()))   # the largest column index is then out of range
    assert ei.value.code == D.E_ARG
    A.close()
