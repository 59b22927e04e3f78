"""Pedigree relationship inverse — host mirror of genmatrix::invertA_henderson
(src/core/a_matrix.cpp:392-439) with ml() inbreeding (a_matrix.cpp:127-186),
over the C-ABI (cb_ainv_*) — and the pedigree file reader that numbers the
animals (parse_ped, src/io/relationship_builder.cpp:83-125)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass
class AinvCSR:
    """A^-1 as CSR rows plus the inbreeding coefficients and d_ii."""
    ptr: np.ndarray   # int64, n + 1
    col: np.ndarray   # int32, nnz
    val: np.ndarray   # float64, nnz
    F: np.ndarray     # float64, n
    dii: np.ndarray   # float64, n

    @property
    def n(self) -> int:
        return self.ptr.size - 1

    def toDense(self) -> np.ndarray:
        A = np.zeros((self.n, self.n))
        rows = np.repeat(np.arange(self.n), np.diff(self.ptr))
        A[rows, self.col] = self.val
        return A


_MISSING_PARENT = {"", "NA", "0", "."}
_HEADER_IDS = {"id", "animal", "ind"}


def read_pedigree(path: str) -> tuple[list[str], np.ndarray, np.ndarray]:
    """parse_ped (relationship_builder.cpp:83-125): whitespace columns
    id, dam, sire taken by position; a line whose first token is id /
    animal / ind (any case) is a header and skipped, as are empty lines;
    "", NA, 0 and . are unknown parents. Animals are numbered by first
    appearance in row order, the row's own id before its dam before its sire,
    so a parent without a row of its own gets a number after its first
    offspring. Returns (ids, dam, sire) with 0-based int32 parents, -1
    unknown; the last row of an id wins. A missing file raises."""
    try:
        f = open(path)
    except OSError:
        raise RuntimeError(f"Cannot open pedigree file: {path}")
    idx: dict[str, int] = {}
    ids: list[str] = []
    rows: list[tuple[str, str, str]] = []
    with f:
        for line in f:
            toks = line.split()
            if not toks:
                continue
            a = toks[0]
            b = toks[1] if len(toks) > 1 else ""
            c = toks[2] if len(toks) > 2 else ""
            if a.lower() in _HEADER_IDS:
                continue
            rows.append((a, b, c))
            for t, is_parent in ((a, False), (b, True), (c, True)):
                if is_parent and t in _MISSING_PARENT:
                    continue
                if t not in idx:
                    idx[t] = len(ids)
                    ids.append(t)
    n = len(ids)
    dam = np.full(n, -1, dtype=np.int32)
    sire = np.full(n, -1, dtype=np.int32)
    for a, b, c in rows:
        dam[idx[a]] = idx.get(b, -1)
        sire[idx[a]] = idx.get(c, -1)
    return ids, dam, sire


def invertA_henderson(dam, sire) -> AinvCSR:
    """Same argument order as the reference: dam first. 0-based parent rows;
    anything outside [0, n) is an unknown parent. A pedigree whose parents
    do not all precede their offspring is solved in a parents-first order
    and mapped back (exact F; the reference's scan is not exact there)."""
    dam = np.ascontiguousarray(dam, dtype=np.int32)
    sire = np.ascontiguousarray(sire, dtype=np.int32)
    if dam.shape != sire.shape:
        raise ValueError("invertA_henderson: dam/sire length mismatch")
    n = dam.size
    h = C.c_void_p()
    _lib.check(_lib.lib().cb_ainv_henderson(n, _lib.ptr(dam), _lib.ptr(sire), C.byref(h)))
    try:
        nnz = C.c_int64()
        _lib.check(_lib.lib().cb_ainv_nnz(h, C.byref(nnz)))
        out = AinvCSR(np.empty(n + 1, np.int64), np.empty(nnz.value, np.int32), np.empty(nnz.value),
                      np.empty(n), np.empty(n))
        _lib.check(_lib.lib().cb_ainv_copy_csr(h, _lib.ptr(out.ptr), _lib.ptr(out.col), _lib.ptr(out.val)))
        _lib.check(_lib.lib().cb_ainv_copy_inbreeding(h, _lib.ptr(out.F), _lib.ptr(out.dii)))
    finally:
        _lib.lib().cb_ainv_destroy(h)
    return out
