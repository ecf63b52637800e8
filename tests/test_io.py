"""On-disk formats of the C-ABI (src/taucs_ccs_io.c): each reader checked
against a matrix whose entries the test states, each writer by a round trip,
and malformed files rejected with NULL rather than a half-built matrix."""
import struct

import numpy as np
import pytest

import taucs_amd as T

SYM = T.TAUCS_SYMMETRIC | T.TAUCS_LOWER | T.TAUCS_DOUBLE


def dense(M):
    n, cp, ri, v = M.arrays()
    D = np.zeros((n, n))
    for j in range(n):
        for p in range(cp[j], cp[j + 1]):
            D[ri[p], j] += v[p]
    return D


def test_binary_round_trip(tmp_path):
    A = T.mesh3d(5)
    f = tmp_path / "a.bin"
    A.write(f, "binary")
    B = T.read_matrix(f, "binary")
    assert B.ptr.contents.flags == A.ptr.contents.flags
    for a, b in zip(A.arrays(), B.arrays()):
        assert np.array_equal(a, b)


def test_binary_layout(tmp_path):
    """nrows, ncols, flags, colptr[ncols+1], rowind[nnz], values[nnz]."""
    f = tmp_path / "m.bin"
    colptr, rowind, vals = [0, 2, 3], [0, 1, 1], [4.0, -1.0, 3.0]
    f.write_bytes(struct.pack("3i", 2, 2, SYM) + struct.pack("3i", *colptr)
                  + struct.pack("3i", *rowind) + struct.pack("3d", *vals))
    M = T.read_matrix(f, "binary")
    assert np.array_equal(dense(M), [[4.0, 0.0], [-1.0, 3.0]])


@pytest.mark.parametrize("cut", [4, 12, 20, 30])
def test_binary_truncated_or_malformed(tmp_path, cut):
    f = tmp_path / "m.bin"
    body = (struct.pack("3i", 2, 2, SYM) + struct.pack("3i", 0, 2, 3) + struct.pack("3i", 0, 1, 1)
            + struct.pack("3d", 4.0, -1.0, 3.0))
    f.write_bytes(body[:cut])
    with pytest.raises(RuntimeError):
        T.read_matrix(f, "binary")
    bad = tmp_path / "bad.bin"   # a row index past the last row
    bad.write_bytes(struct.pack("3i", 2, 2, SYM) + struct.pack("3i", 0, 2, 3)
                    + struct.pack("3i", 0, 5, 1) + struct.pack("3d", 4.0, -1.0, 3.0))
    with pytest.raises(RuntimeError):
        T.read_matrix(bad, "binary")
    dec = tmp_path / "dec.bin"   # decreasing column pointers
    dec.write_bytes(struct.pack("3i", 2, 2, SYM) + struct.pack("3i", 0, 3, 2)
                    + struct.pack("3i", 0, 1, 1) + struct.pack("3d", 4.0, -1.0, 3.0))
    with pytest.raises(RuntimeError):
        T.read_matrix(dec, "binary")


def test_ijv_round_trip_symmetric(tmp_path):
    """A symmetric matrix is written with both triangles ("i j v", 1-based)
    and read back keeping the lower one."""
    A = T.mesh2d(6)
    f = tmp_path / "a.ijv"
    A.write(f, "ijv")
    lines = f.read_text().split("\n")
    entries = [tuple(l.split()) for l in lines if l.strip()]
    n, cp, ri, v = A.arrays()
    offdiag = int(np.sum(ri != np.repeat(np.arange(n), np.diff(cp))))
    assert len(entries) == len(ri) + offdiag
    B = T.read_matrix(f, "ijv", SYM)
    assert np.array_equal(dense(B), dense(A))
    assert np.array_equal(B.to_scipy_full().toarray(), A.to_scipy_full().toarray())


MTX = """%%MatrixMarket matrix coordinate real symmetric
% a comment line
3 3 5
1 1 4.0
2 1 -1.5
2 2 5.0
2 3 -2.0
3 3 6.0
"""


def test_mtx_symmetric_with_banner(tmp_path):
    """An upper entry (2,3) goes to the lower triangle as (3,2)."""
    f = tmp_path / "a.mtx"
    f.write_text(MTX)
    M = T.read_matrix(f, "mtx", SYM)
    assert M.ptr.contents.flags & T.TAUCS_SYMMETRIC
    assert np.array_equal(dense(M), [[4.0, 0, 0], [-1.5, 5.0, 0], [0, -2.0, 6.0]])
    f2 = tmp_path / "b.mtx"      # the same entries without the banner
    f2.write_text("\n".join(MTX.split("\n")[2:]))
    assert np.array_equal(dense(T.read_matrix(f2, "mtx", SYM)), dense(M))


def test_mtx_pattern(tmp_path):
    """TAUCS_PATTERN: a diagonal entry becomes 1 + the largest row index read
    before it, an off-diagonal one -1 (src/taucs_ccs_io.c:704-724)."""
    f = tmp_path / "p.mtx"
    f.write_text("3 3 4\n1 1 9\n3 1 9\n2 2 9\n3 3 9\n")
    M = T.read_matrix(f, "mtx", SYM | T.TAUCS_PATTERN)
    # rows read before each entry: 0, 1, 3, 3
    assert np.array_equal(dense(M), [[1.0, 0, 0], [0, 4.0, 0], [-1.0, 0, 4.0]])


def test_ccs_text(tmp_path):
    """N, colptr[N+1], rowind (0-based), values."""
    f = tmp_path / "a.ccs"
    f.write_text("3\n0 2 3 4\n0 2 1 2\n2.0 -1.0 3.0 5.0\n")
    M = T.read_matrix(f, "ccs", SYM)
    assert np.array_equal(dense(M), [[2.0, 0, 0], [0, 3.0, 0], [-1.0, 0, 5.0]])
    bad = tmp_path / "b.ccs"
    bad.write_text("3\n0 2 3 4\n0 7 1 2\n2.0 -1.0 3.0 5.0\n")
    with pytest.raises(RuntimeError):
        T.read_matrix(bad, "ccs", SYM)


def test_hb_is_not_linked(tmp_path):
    f = tmp_path / "a.rsa"
    f.write_text("not read\n")
    with pytest.raises(RuntimeError):
        T.read_matrix(f, "hb", SYM)


def test_vector_round_trip(tmp_path):
    v = T.random_uniform(37, 5)
    f = tmp_path / "v.bin"
    T.write_vector(f, v)
    assert f.stat().st_size == 37 * 8
    assert np.array_equal(T.read_vector(f, 37), v)
    with pytest.raises(RuntimeError):
        T.read_vector(f, 38)


def test_missing_file(tmp_path):
    for fmt in ("binary", "ijv", "mtx", "ccs"):
        with pytest.raises(RuntimeError):
            T.read_matrix(tmp_path / "none", fmt, SYM)
