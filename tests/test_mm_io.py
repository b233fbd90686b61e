"""CPU tests of the Matrix Market reader (tools/mmio.f mminfo2 / mmread2 and tools/PDRDMM.f:6-90 roles):
the MHD4800 pencil of the reference's example driver (examples/sparse{A,B}.mtx, committed gzip-compressed
under tests/golden) against scipy's independent reader, every supported header variant on small files,
and the error codes.  No GPU needed: the reader is host code."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _gqz():
    import gqz
    return gqz


def test_mhd4800_matches_scipy_reader():
    import gzip
    import scipy.io as sio
    gqz = _gqz()
    for name, symm, nnz in (("mhd4800a.mtx.gz", "general", 102252), ("mhd4800b.mtx.gz", "symmetric", 16160)):
        path = os.path.join(GOLD, name)
        A, info = gqz.mm_read(path)
        assert A.shape == (4800, 4800) and info["rep"] == "coordinate" and info["symm"] == symm
        assert info["nnz"] == nnz
        with gzip.open(path, "rb") as f:
            ref = sio.mmread(f).toarray()
        assert np.array_equal(A, ref)


def _write(tmp_path, text):
    p = tmp_path / "m.mtx"
    p.write_text(text)
    return str(p)


@pytest.mark.parametrize("text,expect", [
    ("%%MatrixMarket matrix coordinate real general\n% c\n3 3 2\n1 1 2.5\n3 2 -1\n",
     [[2.5, 0, 0], [0, 0, 0], [0, -1, 0]]),
    ("%%MatrixMarket matrix coordinate real symmetric\n3 3 2\n1 1 1\n3 1 4\n",
     [[1, 0, 4], [0, 0, 0], [4, 0, 0]]),
    ("%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 3\n", [[0, -3], [3, 0]]),
    ("%%MatrixMarket matrix coordinate pattern general\n2 3 2\n1 3\n2 1\n", [[0, 0, 1], [1, 0, 0]]),
    ("%%MatrixMarket matrix coordinate integer general\n2 2 1\n2 2 7\n", [[0, 0], [0, 7]]),
    ("%%MatrixMarket matrix array real general\n2 3\n1\n2\n3\n4\n5\n6\n", [[1, 3, 5], [2, 4, 6]]),
    ("%%MatrixMarket matrix array real symmetric\n2 2\n1\n2\n3\n", [[1, 2], [2, 3]]),
])
def test_header_variants(tmp_path, text, expect):
    gqz = _gqz()
    A, _ = gqz.mm_read(_write(tmp_path, text))
    assert np.array_equal(A, np.array(expect, dtype=float))


@pytest.mark.parametrize("text,code", [
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", -3),
    ("not a matrix market file\n", -2),
    ("%%MatrixMarket matrix coordinate real general\n2 2\n", -4),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n", -5),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", -6),
])
def test_error_codes(tmp_path, text, code):
    import ctypes as C
    gqz = _gqz()
    L = gqz.lib()
    p = _write(tmp_path, text).encode()
    m, n, nnz = C.c_int(0), C.c_int(0), C.c_long(0)
    r = L.gqz_mm_info(p, None, None, None, C.byref(m), C.byref(n), C.byref(nnz))
    if r == 0:
        A = np.zeros((max(1, m.value), max(1, n.value)), order="F")
        r = L.gqz_mm_read_dense(p, m.value, n.value, A.ctypes.data_as(C.c_void_p), max(1, m.value))
    assert r == code
    assert L.gqz_mm_read_dense(b"/nonexistent/file.mtx", 1, 1, None, 1) == -1
