"""Relationship-matrix files — host mirror of write_named_matrix
(src/io/relationship_builder.cpp:43-53), write_matrix_txt / write_matrix_bin
(src/core/matrix_io.cpp:464-479) and read/write_upper_packed_binary
(src/core/upper_packed_binary.cpp), over the C-ABI (cb_matrix_write,
cb_upper_packed_read, cb_grm_write)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _lib

FORMATS = {"txt": _lib.CB_MATRIX_TXT, "bin": _lib.CB_MATRIX_BIN}


def write_matrix(path: str, M: np.ndarray, fmt: str = "bin") -> None:
    """A host symmetric matrix as write_matrix_txt ("txt") or
    write_upper_packed_binary ("bin") writes it."""
    A = np.ascontiguousarray(M, dtype=np.float64)
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ValueError("upper-packed binary output requires a square matrix: " + str(path))
    _lib.check(_lib.lib().cb_matrix_write(str(path).encode(), A.shape[0], _lib.ptr(A), A.shape[1], FORMATS[fmt]))


def write_upper_packed_binary(path: str, M: np.ndarray) -> None:
    write_matrix(path, M, "bin")


def read_upper_packed_binary(path: str) -> np.ndarray:
    """read_upper_packed_binary: the full symmetric matrix."""
    n = C.c_int64()
    _lib.check(_lib.lib().cb_upper_packed_read(str(path).encode(), C.byref(n), None, 0))
    out = np.empty((n.value, n.value))
    _lib.check(_lib.lib().cb_upper_packed_read(str(path).encode(), C.byref(n), _lib.ptr(out), n.value))
    return out


def read_matrix_txt(path: str, n: int) -> np.ndarray:
    """Triplets "i j value" (1-based) as write_matrix_txt writes them,
    expanded symmetrically into an n x n matrix."""
    t = np.loadtxt(path, ndmin=2)
    M = np.zeros((n, n))
    if t.size:
        i = t[:, 0].astype(np.int64) - 1
        j = t[:, 1].astype(np.int64) - 1
        M[i, j] = t[:, 2]
        M[j, i] = t[:, 2]
    return M


def write_ids(path: str, ids) -> None:
    with open(path, "w") as f:
        for x in ids:
            f.write(f"{x}\n")


def write_named_matrix(out_prefix: str, stem: str, ids, fmt: str = "bin", matrix: np.ndarray | None = None,
                       grm=None) -> tuple[str, str]:
    """write_named_matrix: <prefix>.<stem>.txt or .bin plus <prefix>.<stem>.id.
    The matrix comes from a host array or straight from a GrmBuilder's FP64
    G on the device. Returns (matrix path, id path)."""
    if (matrix is None) == (grm is None):
        raise ValueError("write_named_matrix: give exactly one of matrix, grm")
    mpath = f"{out_prefix}.{stem}.{fmt}"
    ipath = f"{out_prefix}.{stem}.id"
    Path(mpath).parent.mkdir(parents=True, exist_ok=True)
    if grm is not None:
        grm.writeMatrix(mpath, fmt)
    else:
        write_matrix(mpath, matrix, fmt)
    write_ids(ipath, ids)
    return mpath, ipath
