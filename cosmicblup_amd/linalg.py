"""Relationship-matrix inverse — host mirror of invert_psd_matrix
(src/core/matrix_io.cpp:488-501) over the C-ABI (cb_invert_psd*)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def invert_psd_matrix(M: np.ndarray, ridge: float = 0.0, ctx: _lib.Context | None = None) -> np.ndarray:
    """Eigen-bent inverse of a symmetric PSD matrix (host arrays)."""
    ctx = ctx or _lib.default_context()
    M = np.ascontiguousarray(M, dtype=np.float64)
    n = M.shape[0]
    if M.shape != (n, n):
        raise ValueError("invert_psd_matrix: square matrix required")
    out = np.empty_like(M)
    _lib.check(_lib.lib().cb_invert_psd(ctx.handle, _lib.ptr(M), n, n, float(ridge), _lib.ptr(out), n, None, None))
    return out


def invert_psd_device(dM: int, ldm: int, n: int, dout: int, ldo: int, ridge: float = 0.0,
                      ctx: _lib.Context | None = None, eigenvalues: bool = False):
    """Same on device memory (may alias). With eigenvalues=True returns the
    smallest and largest eigenvalue of M + ridge I before the floor (the
    library then takes the full eigensolve); otherwise (None, None), and the
    library inverts by Cholesky where the floor allows it."""
    ctx = ctx or _lib.default_context()
    if not eigenvalues:
        _lib.check(_lib.lib().cb_invert_psd_device(ctx.handle, dM, ldm, n, float(ridge), dout, ldo, None, None))
        return None, None
    lo, hi = C.c_double(), C.c_double()
    _lib.check(_lib.lib().cb_invert_psd_device(ctx.handle, dM, ldm, n, float(ridge), dout, ldo,
                                               C.byref(lo), C.byref(hi)))
    return lo.value, hi.value
