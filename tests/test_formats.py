"""The reference's matrix and vector file formats (MatrixCL::exportMTX / out / in,
VecDescBaseCL::Write / Read) written by libdhip, byte for byte against the same
statements run through C++ iostreams (oracle/format_oracle.cpp).  Host memory only:
no GPU needed."""
import os

import numpy as np
import pytest

import oracle as O
from drops_amd import dhip as D


def random_csr(n, per_row, seed, ncols=None, empty_every=7):
    rng = np.random.default_rng(seed)
    ncols = n if ncols is None else ncols
    rows = []
    for r in range(n):
        k = 0 if (empty_every and r % empty_every == 3) else int(rng.integers(1, 2 * per_row))
        rows.append(np.sort(rng.choice(ncols, size=min(k, ncols), replace=False)))
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum([len(c) for c in rows])
    ci = np.concatenate(rows) if rp[-1] else np.zeros(0, np.int64)
    val = rng.standard_normal(ci.size) * 10.0 ** rng.integers(-30, 30, ci.size)
    return rp, ci, val


SPECIAL = np.array([0.0, -0.0, 1.0, -1.0, 0.1, 1e-310, -4.9e-324, 1.7976931348623157e308, 123456789012345.0,
                    2.0 ** -1074, 1.0 / 3.0, 1e15, 1e16, 9.999999999999999e22, 0.5e-5, 1e-5])


def _same_bytes(a, b):
    with open(a, "rb") as fa, open(b, "rb") as fb:
        x, y = fa.read(), fb.read()
    assert x == y, (len(x), len(y), next((i for i in range(min(len(x), len(y))) if x[i] != y[i]), None))


@pytest.mark.parametrize("fmt", ["mtx", "plain"])
@pytest.mark.parametrize("case", ["random", "special", "large", "empty", "empty_rows", "rect"])
def test_matrix_bytes_match_iostreams(tmp_path, fmt, case):
    ncols = None
    if case == "random":
        rp, ci, v = random_csr(500, 6, 1)
    elif case == "special":
        rp = np.arange(SPECIAL.size + 1)
        ci = np.arange(SPECIAL.size)
        v = SPECIAL
    elif case == "large":  # several formatting threads
        rp, ci, v = random_csr(60000, 6, 2)
    elif case == "empty":
        rp, ci, v = np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0)
    elif case == "empty_rows":
        rp, ci, v = np.zeros(6, np.int64), np.zeros(0, np.int64), np.zeros(0)
    else:
        rp, ci, v = random_csr(40, 3, 3, ncols=97)
        ncols = 97
    ours, ref = str(tmp_path / "ours"), str(tmp_path / "ref")
    D.mtx_write(ours, rp, ci, v, ncols=ncols, fmt=fmt)
    O.fmt_matrix(ref, rp, ci, v, ncols=ncols, fmt=fmt)
    _same_bytes(ours, ref)


def test_matrix_nonfinite_values(tmp_path):
    v = np.array([np.inf, -np.inf, np.nan])
    rp, ci = np.arange(4), np.arange(3)
    for fmt in ("mtx", "plain"):
        ours, ref = str(tmp_path / ("o" + fmt)), str(tmp_path / ("r" + fmt))
        D.mtx_write(ours, rp, ci, v, fmt=fmt)
        O.fmt_matrix(ref, rp, ci, v, fmt=fmt)
        _same_bytes(ours, ref)


def test_plain_read_matches_in(tmp_path):
    rp, ci, v = random_csr(3000, 5, 4)
    path = str(tmp_path / "m.txt")
    O.fmt_matrix(path, rp, ci, v, fmt="plain")
    ours = D.mtx_read(path, "plain")
    ref = O.read_plain(path)
    assert ours["shape"] == ref["shape"] == (3000, 3000)
    assert np.array_equal(ours["rowptr"], ref["rowptr"]) and np.array_equal(ours["colind"], ref["colind"])
    assert np.array_equal(ours["val"], ref["val"])
    # 6 significant digits, as the reference's out() writes them
    assert np.allclose(ours["val"], v, rtol=5e-6, atol=0)


def test_mtx_round_trip_and_unsorted_input(tmp_path):
    rp, ci, v = random_csr(2000, 8, 5)
    path = str(tmp_path / "a.mtx")
    D.mtx_write(path, rp, ci, v)
    back = D.mtx_read(path)
    assert back["shape"] == (2000, 2000)
    assert np.array_equal(back["rowptr"], rp) and np.array_equal(back["colind"], ci)
    # %.15e keeps 16 significant digits
    assert np.all(np.abs(back["val"] - v) <= 1e-15 * np.abs(v))
    # entries in any order, repeated (row, col) pairs summed (values exact in binary)
    lines = ["%%MatrixMarket matrix coordinate real general", "% a comment", "3 4 6",
             "3 1 1.5", "1 4 2", "1 2 0.25", "3 1 0.5", "2 2 -1", "1 2 4"]
    p2 = str(tmp_path / "b.mtx")
    with open(p2, "w") as f:
        f.write("\n".join(lines) + "\n")
    m = D.mtx_read(p2)
    assert m["shape"] == (3, 4)
    assert m["rowptr"].tolist() == [0, 2, 3, 4]
    assert m["colind"].tolist() == [1, 3, 1, 0]
    assert m["val"].tolist() == [4.25, 2.0, -1.0, 2.0]


def test_vector_bytes_and_read(tmp_path):
    rng = np.random.default_rng(6)
    for x in (np.concatenate([SPECIAL, rng.standard_normal(70000)]), np.zeros(0)):
        for binary in (True, False):
            ours, ref = str(tmp_path / "ov"), str(tmp_path / "rv")
            D.vec_write(ours, x, binary=binary)
            O.fmt_vector(ref, x, binary=binary)
            _same_bytes(ours, ref)
            assert D.vec_count(ours, binary) == x.size
            got = D.vec_read(ours, x.size, binary=binary)
            if binary:
                assert np.array_equal(got.view(np.uint64), x.view(np.uint64))
                assert np.array_equal(O.read_vector(ref, x.size).view(np.uint64), x.view(np.uint64))
            else:
                # 15 significant digits; subnormals keep fewer bits
                assert np.all(np.abs(got - x) <= 1e-14 * np.abs(x) + 1e-320)
    # text reads against operator>>, which stops at subnormal values (failbit on underflow):
    # compared on normal numbers only
    x = rng.standard_normal(5000) * 10.0 ** rng.integers(-300, 300, 5000)
    path = str(tmp_path / "t")
    D.vec_write(path, x, binary=False)
    assert np.array_equal(D.vec_read(path, x.size, binary=False), O.read_vector(path, x.size, binary=False))


def test_format_errors(tmp_path):
    x = np.arange(5.0)
    path = str(tmp_path / "v.bin")
    D.vec_write(path, x)
    with pytest.raises(D.DhipError) as e:
        D.vec_read(path, 4)
    assert e.value.code == D.E_ARG and "Number of Unknowns" in str(e.value)
    with pytest.raises(O.OracleError) as eo:
        O.read_vector(path, 4)
    assert eo.value.code == 1
    with pytest.raises(D.DhipError) as e:
        D.mtx_read(str(tmp_path / "missing.mtx"))
    assert e.value.code == D.E_IO
    bad = str(tmp_path / "bad.mtx")
    with open(bad, "w") as f:
        f.write("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    with pytest.raises(D.DhipError) as e:
        D.mtx_read(bad)
    assert e.value.code == D.E_IO
    with open(bad, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n")
    with pytest.raises(D.DhipError) as e:
        D.mtx_read(bad)
    assert e.value.code == D.E_IO
    with pytest.raises(D.DhipError):
        D.mtx_write(os.path.join(str(tmp_path), "no", "such", "dir.mtx"), np.array([0]), np.zeros(0), np.zeros(0))


def test_buffer_variants_match_files(tmp_path):
    """dhip_mtx_format / dhip_vec_format give the file's bytes within the stated bound;
    dhip_mtx_parse / dhip_vec_parse read them back as the file readers do; a short buffer
    is refused."""
    rp, ci, v = random_csr(20000, 6, 7)
    for fmt in ("mtx", "plain"):
        path = str(tmp_path / ("m" + fmt))
        D.mtx_write(path, rp, ci, v, fmt=fmt)
        data = D.mtx_format(rp, ci, v, fmt=fmt)
        with open(path, "rb") as f:
            assert f.read() == data
        a, b = D.mtx_parse(data, fmt), D.mtx_read(path, fmt)
        for k in ("rowptr", "colind", "val"):
            assert np.array_equal(a[k], b[k])
    x = np.concatenate([SPECIAL, np.random.default_rng(8).standard_normal(30000)])
    for binary in (True, False):
        path = str(tmp_path / "v")
        D.vec_write(path, x, binary=binary)
        data = D.vec_format(x, binary=binary)
        with open(path, "rb") as f:
            assert f.read() == data
        assert np.array_equal(D.vec_parse(data, x.size, binary), D.vec_read(path, x.size, binary))
    L = D.load()
    import ctypes as C
    buf = C.create_string_buffer(100)
    ln = C.c_int64()
    a = np.ascontiguousarray(x)
    rc = L.dhip_vec_format(None, a.ctypes.data, a.size, D.HOST, 0, C.cast(buf, C.c_void_p), 100, C.byref(ln))
    assert rc == D.E_CAPACITY


def test_size_bounds_hold_for_extreme_values():
    """the upper bound a size query returns covers the longest items (most negative
    exponents, largest indices a test can afford)"""
    n = 3
    rp = np.array([0, 2, 3, 4])
    ci = np.array([2, 1, 0, 2])
    v = np.array([-1.7976931348623157e308, -2.2250738585072014e-308, -4.9e-324, -1.2345678901234567e-300])
    L = D.load()
    import ctypes as C
    for fmt in (D.MTX_MARKET, D.MTX_PLAIN):
        cap = C.c_int64()
        r32, c32 = rp.astype(np.int32), ci.astype(np.int32)
        args = (fmt, n, n, r32.ctypes.data_as(C.POINTER(C.c_int32)), c32.ctypes.data_as(C.POINTER(C.c_int32)),
                v.ctypes.data_as(C.POINTER(C.c_double)))
        assert L.dhip_mtx_format(*args, None, 0, C.byref(cap)) == 0
        assert len(D.mtx_format(rp, ci, v, fmt=fmt)) <= cap.value
    assert len(D.vec_format(v, binary=False)) <= 128 + 24 * v.size
