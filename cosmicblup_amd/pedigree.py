"""Pedigree relationship inverse — host mirror of genmatrix::invertA_henderson
(src/core/a_matrix.cpp:392-439) with ml() inbreeding (a_matrix.cpp:127-186),
over the C-ABI (cb_ainv_*)."""
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


def invertA_henderson(dam, sire) -> AinvCSR:
    """Same argument order as the reference: dam first. 0-based parent rows;
    anything outside [0, n) is an unknown parent. Parents must precede
    offspring."""
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
