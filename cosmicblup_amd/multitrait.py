"""Multi-trait GBLUP — host mirrors of MvEigenVce (src/core/eigen_vce.cpp:
106-262) and GSSolve::solveMultiTrait (src/solve/mixed_model_solver.cpp:
425-524) over the C-ABI (cb_mv_*). Device buffers are torch tensors
(plumbing only)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


def _device(M):
    import torch

    if torch.is_tensor(M):
        return M.contiguous()
    return torch.from_numpy(np.ascontiguousarray(M, dtype=np.float64)).cuda()


class MvEigenVce:
    """prepare(Y, X, G) then runNullModel(); getVg() / getVe()."""

    @dataclass
    class Options:
        use_projection: bool = True
        max_iter: int = 500
        tol: float = 1e-4

    def __init__(self, options: "MvEigenVce.Options | None" = None, ctx: _lib.Context | None = None):
        self.options = options or MvEigenVce.Options()
        if not self.options.use_projection:
            raise ValueError("MvEigenVce: only the projected form (the reference's multi-trait path) is built")
        self.ctx = ctx or _lib.default_context()
        self.Vg = self.Ve = None
        self.iterations = 0
        self.converged = False

    def prepare(self, Y, X, G) -> None:
        self.Y = np.asfortranarray(np.asarray(Y, dtype=np.float64))
        self.X = np.asfortranarray(np.asarray(X, dtype=np.float64))
        self.G = _device(G)

    def runNullModel(self) -> bool:
        import torch

        n, t = self.Y.shape
        c = self.X.shape[1]
        o = _lib.MvOptions()
        _lib.lib().cb_mv_options_default(C.byref(o))
        o.max_iter, o.tol = self.options.max_iter, self.options.tol
        Vg = np.empty((t, t))
        Ve = np.empty((t, t))
        it = C.c_int64()
        conv = C.c_int()
        Yf = np.ascontiguousarray(self.Y.T)  # column-major n x t
        Xf = np.ascontiguousarray(self.X.T)
        torch.cuda.current_stream().synchronize()  # torch-made inputs ready (not other streams)
        _lib.check(_lib.lib().cb_mv_vce(self.ctx.handle, n, t, c, _lib.ptr(Yf), _lib.ptr(Xf), self.G.data_ptr(),
                                        n, C.byref(o), _lib.ptr(Vg), _lib.ptr(Ve), C.byref(it), C.byref(conv)))
        self.Vg, self.Ve = Vg, Ve
        self.iterations, self.converged = it.value, bool(conv.value)
        return self.converged

    def getVg(self) -> np.ndarray:
        return self.Vg

    def getVe(self) -> np.ndarray:
        return self.Ve


def solveMultiTrait(Y, X, Ginv, Vg, Ve, ctx: _lib.Context | None = None):
    """Solution of the Kronecker MME of GSSolve::solveMultiTrait
    (mixed_model_solver.cpp:425-524) by the canonical transformation and t
    single-trait PCG solves on the device. Returns (beta c x t, u n x t)."""
    import torch

    ctx = ctx or _lib.default_context()
    Y = np.asarray(Y, dtype=np.float64)
    X = np.asarray(X, dtype=np.float64)
    n, t = Y.shape
    c = X.shape[1]
    dG = _device(Ginv)
    beta = np.empty((t, c))
    u = np.empty((t, n))
    Yf = np.ascontiguousarray(Y.T)
    Xf = np.ascontiguousarray(X.T)
    Vg = np.ascontiguousarray(Vg, dtype=np.float64)
    Ve = np.ascontiguousarray(Ve, dtype=np.float64)
    torch.cuda.current_stream().synchronize()  # torch-made inputs ready (not other streams)
    _lib.check(_lib.lib().cb_mv_solve(ctx.handle, n, t, c, _lib.ptr(Yf), _lib.ptr(Xf), dG.data_ptr(), n, _lib.ptr(Vg),
                                      _lib.ptr(Ve), _lib.ptr(beta), _lib.ptr(u)))
    return beta.T, u.T
