"""MatrixMarket reading and writing (tg_mm_info / tg_mm_read / tg_mm_write), the host half
of MatrixCSR::LoadFromMatrixMarket / WriteToMatrixMarket (MatrixCSR/MatrixMarket.hpp:3-306).
Host-only entry points: these run without a GPU.  The parse is checked against an
independent reading of the same text (numpy / scipy.io.mmread) and by exact round trips."""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from tensors_amd import _capi as capi


def info(path):
    m, n, cnt, field, symm = C.c_int32(0), C.c_int32(0), C.c_int64(0), C.c_int32(0), C.c_int32(0)
    rc = capi.lib().tg_mm_info(os.fsencode(path), C.byref(m), C.byref(n), C.byref(cnt), C.byref(field),
                               C.byref(symm))
    return rc, (m.value, n.value, cnt.value, field.value, symm.value)


def read(path, threads=0):
    rc, (m, n, k, field, symm) = info(path)
    assert rc == 0, capi.lib().tg_last_error()
    I, J, V = np.empty(max(k, 1), np.int32), np.empty(max(k, 1), np.int32), np.empty(max(k, 1), np.float64)
    rc = capi.lib().tg_mm_read(os.fsencode(path), k, I.ctypes.data, J.ctypes.data, V.ctypes.data, threads)
    return rc, m, n, field, symm, I[:k], J[:k], V[:k]


def write(path, m, n, rp, ci, vals, threads=0):
    rp = np.ascontiguousarray(rp)
    ci = np.ascontiguousarray(ci, np.int32)
    tv = capi.F64 if vals is None or vals.dtype == np.float64 else capi.F32
    return capi.lib().tg_mm_write(os.fsencode(path), m, n, rp.ctypes.data, capi.I64 if rp.dtype == np.int64 else capi.I32,
                                  ci.ctypes.data if ci.size else None, None if vals is None else vals.ctypes.data,
                                  tv, threads)


def random_csr(m, n, density, seed, dtype=np.float64):
    A = sp.random(m, n, density=density, format="csr", random_state=seed, dtype=np.float64)
    A.data = np.random.default_rng(seed).normal(size=A.nnz) * 10.0 ** np.random.default_rng(seed + 1).integers(
        -30, 30, A.nnz)
    A.sort_indices()
    return A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.astype(dtype)


@pytest.mark.parametrize("threads", [1, 3, 0])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_write_read_round_trip_is_exact(tmp_path, threads, dtype):
    m, n = 700, 500
    rp, ci, val = random_csr(m, n, 0.02, 7, dtype)
    f = tmp_path / "a.mtx"
    assert write(f, m, n, rp, ci, val, threads) == 0
    rc, m2, n2, field, symm, I, J, V = read(f, threads)
    assert rc == 0 and (m2, n2, field, symm) == (m, n, 0, 0)
    rows = np.repeat(np.arange(m), np.diff(rp))
    assert np.array_equal(I, rows) and np.array_equal(J, ci)
    assert np.array_equal(V.astype(dtype), val)  # %.17E / %.8E carry every bit
    # an independent reader sees the same matrix
    S = scipy.io.mmread(str(f)).tocsr()
    S.sort_indices()
    assert np.array_equal(S.indptr, rp) and np.array_equal(S.indices, ci)
    assert np.array_equal(S.data.astype(dtype), val)


def test_written_text_format(tmp_path):
    rp = np.array([0, 2, 2, 3], np.int32)
    ci = np.array([0, 2, 1], np.int32)
    val = np.array([1.5, -2.0, 1e-300])
    f = tmp_path / "b.mtx"
    assert write(f, 3, 3, rp, ci, val, 2) == 0
    lines = f.read_text().splitlines()
    assert lines[0] == "%%MatrixMarket matrix coordinate real general"
    assert lines[1] == "3 3 3"
    assert lines[2:] == ["1 1 1.50000000000000000E+00", "1 3 -2.00000000000000000E+00",
                         "3 2 1.00000000000000003E-300"]
    # pattern (no values)
    assert write(f, 3, 3, rp, ci, None) == 0
    lines = f.read_text().splitlines()
    assert lines[0] == "%%MatrixMarket matrix coordinate pattern general" and lines[2:] == ["1 1", "1 3", "3 2"]


def test_reader_accepts_what_the_format_allows(tmp_path):
    """Mixed-case banner, comments, blank lines, tabs, '+' signs, exponents, CRLF."""
    text = ("%%MatrixMarket Matrix Coordinate Real General\r\n"
            "% a comment\n"
            "%\n"
            "\n"
            "4 5 6\n"
            "1 1 +1.0\n"
            "\t2  3\t-2.5e-3\n"
            "\n"
            "4 5 7\r\n"
            "3 1 1E+10\n"
            "2 3 0.25\n"
            "1 5 -0\n")
    f = tmp_path / "c.mtx"
    f.write_text(text)
    rc, m, n, field, symm, I, J, V = read(f, 4)
    assert rc == 0 and (m, n, field, symm) == (4, 5, 0, 0)
    assert I.tolist() == [0, 1, 3, 2, 1, 0] and J.tolist() == [0, 2, 4, 0, 2, 4]
    assert V.tolist() == [1.0, -2.5e-3, 7.0, 1e10, 0.25, 0.0]


def test_pattern_integer_and_symmetric_headers(tmp_path):
    f = tmp_path / "p.mtx"
    f.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n3 3 3\n1 1\n2 1\n3 2\n")
    rc, m, n, field, symm, I, J, V = read(f)
    assert rc == 0 and (field, symm) == (2, 1) and V.tolist() == [1.0, 1.0, 1.0]
    assert I.tolist() == [0, 1, 2] and J.tolist() == [0, 0, 1]
    f.write_text("%%MatrixMarket matrix coordinate integer general\n2 2 2\n1 2 -7\n2 2 3\n")
    rc, m, n, field, symm, I, J, V = read(f)
    assert rc == 0 and (field, symm) == (1, 0) and V.tolist() == [-7.0, 3.0]


@pytest.mark.parametrize("text,code", [
    ("%%NotMarket matrix coordinate real general\n1 1 0\n", capi.ERR_INVALID),
    ("%%MatrixMarket tensor coordinate real general\n1 1 0\n", capi.ERR_INVALID),
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", capi.ERR_INVALID),
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", capi.ERR_DTYPE),
    ("%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1\n", capi.ERR_INVALID),
    ("%%MatrixMarket matrix coordinate real hermitian\n2 2 1\n2 1 1\n", capi.ERR_INVALID),
    ("%%MatrixMarket matrix coordinate real upper\n2 2 1\n2 1 1\n", capi.ERR_INVALID),
    ("%%MatrixMarket matrix coordinate real general\n-2 2 1\n1 1 1\n", capi.ERR_INVALID),
])
def test_header_errors(tmp_path, text, code):
    f = tmp_path / "e.mtx"
    f.write_text(text)
    rc, _ = info(f)
    assert rc == code


@pytest.mark.parametrize("body", [
    "3 3 2\n1 1 1.0\n",            # fewer entries than announced
    "3 3 1\n4 1 1.0\n",            # row index out of range
    "3 3 1\n1 0 1.0\n",            # column index 0 (the format is one-based)
    "3 3 1\n1 1 x\n",              # malformed value
])
def test_entry_errors(tmp_path, body):
    f = tmp_path / "e.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n" + body)
    rc = read(f)[0]
    assert rc == capi.ERR_INVALID
    assert b"tg_mm_read" in capi.lib().tg_last_error()


def test_missing_file_and_extra_entries(tmp_path):
    assert info(tmp_path / "none.mtx")[0] == capi.ERR_INVALID
    f = tmp_path / "x.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 5\n2 2 6\n")
    rc, m, n, field, symm, I, J, V = read(f)
    assert rc == 0 and I.tolist() == [0] and V.tolist() == [5.0]  # the announced count is read


def test_large_file_parallel_parse_matches_numpy(tmp_path):
    """A 200k-entry file split over many threads parses exactly as numpy reads the text."""
    rng = np.random.default_rng(3)
    k = 200_000
    I = rng.integers(1, 5001, k)
    J = rng.integers(1, 3001, k)
    V = rng.normal(size=k)
    f = tmp_path / "big.mtx"
    with open(f, "w") as fh:
        fh.write("%%MatrixMarket matrix coordinate real general\n" + f"5000 3000 {k}\n")
        np.savetxt(fh, np.column_stack([I, J, V]), fmt=["%d", "%d", "%.17g"])
    for threads in (1, 7, 0):
        rc, m, n, field, symm, I2, J2, V2 = read(f, threads)
        assert rc == 0
        assert np.array_equal(I2, I - 1) and np.array_equal(J2, J - 1) and np.array_equal(V2, V)
