"""Relationship-matrix files (write_matrix_txt, write_matrix_bin,
upper_packed_binary.cpp): the byte layout of COSMIC_UPPER_PACKED written by
hand from its header definition (upper_packed_binary.h:5-18), the
reference's read checks, and the text triplet format. Host code only."""
import struct

import numpy as np
import pytest

from cosmicblup_amd import CosmicError, matrixio


def sym(n, seed=0):
    A = np.random.default_rng(seed).standard_normal((n, n))
    return A + A.T


def packed_bytes(A):
    """The format from its definition: magic[32], i32 version 1, i64 n,
    i32 index base 1, i32 storage 1, i32 dtype 1, i32 endian 1, then the
    upper triangle column by column as little-endian doubles."""
    n = A.shape[0]
    head = b"COSMIC_UPPER_PACKED".ljust(32, b"\0") + struct.pack("<iqiiii", 1, n, 1, 1, 1, 1)
    body = b"".join(struct.pack("<d", A[i, j]) for j in range(n) for i in range(j + 1))
    return head + body


def test_upper_packed_layout_and_round_trip(cb, tmp_path):
    A = sym(7)
    p = tmp_path / "g.bin"
    matrixio.write_upper_packed_binary(str(p), A)
    assert p.read_bytes() == packed_bytes(A)
    np.testing.assert_array_equal(matrixio.read_upper_packed_binary(str(p)), A)
    # a hand-written file reads back too
    B = sym(4, seed=1)
    q = tmp_path / "h.bin"
    q.write_bytes(packed_bytes(B))
    np.testing.assert_array_equal(matrixio.read_upper_packed_binary(str(q)), B)


@pytest.mark.parametrize("damage,msg", [
    (lambda b: b"X" + b[1:], "unsupported binary matrix header"),
    (lambda b: b[:32] + struct.pack("<i", 2) + b[36:], "unsupported upper-packed binary matrix metadata"),
    (lambda b: b[:36] + struct.pack("<q", 0) + b[44:], "invalid upper-packed binary matrix dimension"),
    (lambda b: b[:-3], "truncated upper-packed binary matrix file"),
    (lambda b: b + b"\0", "trailing bytes"),
])
def test_upper_packed_reader_checks(cb, tmp_path, damage, msg):
    p = tmp_path / "bad.bin"
    p.write_bytes(damage(packed_bytes(sym(3))))
    with pytest.raises(CosmicError, match=msg):
        matrixio.read_upper_packed_binary(str(p))


def test_reader_missing_file(cb, tmp_path):
    with pytest.raises(CosmicError, match="cannot read upper-packed binary matrix file"):
        matrixio.read_upper_packed_binary(str(tmp_path / "absent.bin"))


def test_text_triplets(cb, tmp_path):
    """write_matrix_txt: i<TAB>j<TAB>value for j <= i, 1-based, values as a
    C++ stream prints a double by default (6 significant digits, %g)."""
    A = sym(5)
    A[2, 1] = A[1, 2] = 1.0e-5
    A[3, 0] = A[0, 3] = 123456789.0
    p = tmp_path / "g.txt"
    matrixio.write_matrix(str(p), A, "txt")
    lines = p.read_text().splitlines()
    expect = [f"{i + 1}\t{j + 1}\t{A[i, j]:g}" for i in range(5) for j in range(i + 1)]
    assert lines == expect
    assert "3\t2\t1e-05" in lines and "4\t1\t1.23457e+08" in lines
    back = matrixio.read_matrix_txt(str(p), 5)
    np.testing.assert_allclose(back, A, rtol=1e-5)


def test_named_matrix_files(cb, tmp_path):
    """write_named_matrix: <prefix>.<stem>.{txt|bin} and <prefix>.<stem>.id."""
    A = sym(3)
    ids = ["s1", "s2", "s3"]
    m, i = matrixio.write_named_matrix(str(tmp_path / "out/run"), "GA", ids, "bin", matrix=A)
    assert m.endswith("run.GA.bin") and i.endswith("run.GA.id")
    assert open(i).read() == "s1\ns2\ns3\n"
    assert open(m, "rb").read()[:19] == b"COSMIC_UPPER_PACKED"
    with pytest.raises(CosmicError, match="Cannot write matrix file"):
        matrixio.write_matrix(str(tmp_path / "no/such/dir/g.txt"), A, "txt")
