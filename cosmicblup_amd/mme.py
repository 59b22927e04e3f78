"""Mixed-model equations and REML — host mirrors of the reference's solver
interfaces over the C-ABI (cb_mme_*, cb_reml_*).

  MixedModel       MMELHSBuilder + the relationship inverse of RandomComponent
                   (src/core/mme_builder.{h,cpp}, src/core/matrix_adapter.h)
  MixedModel.solve PCGSolver / MatrixFreePCG (src/core/pcg_solver.h,
                   src/core/matrix_free_pcg.h)
  AIREML           VCESolver / AI_REML (src/core/vce.h:223-547): solve(),
                   getVarE(), getVarsU(), getSolution(), history
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .design import FixedDesignG


def solve_options(precision=_lib.CB_PREC_MIXED, precond=_lib.CB_PRECOND_BLOCK_JACOBI,
                  stop_rule=_lib.CB_STOP_RELATIVE, tol=1e-10, max_iter=10000,
                  batch_rhs=True, stall_tol=0.0) -> _lib.SolveOptions:
    """stall_tol 0 (the library default): a mixed-precision solve that stalls
    at the FP64 rounding floor above tol raises NotConverged."""
    o = _lib.SolveOptions()
    _lib.lib().cb_solve_options_default(C.byref(o))
    o.precision, o.precond, o.stop_rule = precision, precond, stop_rule
    o.tol, o.max_iter, o.batch_rhs = tol, max_iter, int(batch_rhs)
    o.stall_tol = stall_tol
    return o


def reml_options(algorithm=_lib.CB_VCE_AI, max_iter=20, tol=1e-4, n_probes=64, seed=42,
                 trace_exact=False, solve: _lib.SolveOptions | None = None,
                 block_cols: int = 0, stall_tol: float = 1e-8, mc_samples: int = 100,
                 adaptive: bool = False) -> _lib.RemlOptions:
    """stall_tol (default 1e-8, as cb_reml_options_default) is applied to the
    REML solves, also to a `solve` given here. algorithm CB_VCE_AUTO takes
    the reference's size-based branch (AIREML.branch() reports it)."""
    o = _lib.RemlOptions()
    _lib.lib().cb_reml_options_default(C.byref(o))
    o.algorithm, o.max_iter, o.tol, o.n_probes, o.seed = algorithm, max_iter, tol, n_probes, seed
    o.trace_exact = int(trace_exact)
    o.block_cols = block_cols
    o.mc_samples, o.adaptive = mc_samples, int(adaptive)
    if solve is not None:
        o.solve = solve
    o.solve.stall_tol = stall_tol
    return o


class MixedModel:
    """One random component: y = X b + Z u + e, Var(u) = sigma_u^2 Q^-1."""

    def __init__(self, fd: FixedDesignG, aid: np.ndarray, y: np.ndarray, q: int,
                 ctx: _lib.Context | None = None):
        self.ctx = ctx or _lib.default_context()
        self.fd = fd
        self.n = int(y.size)
        self.p = int(fd.p)
        self.q = int(q)
        self._keep = [np.ascontiguousarray(a) for a in (fd.ptr, fd.col, fd.val)]
        aid = np.ascontiguousarray(aid, dtype=np.int32)
        y = np.ascontiguousarray(y, dtype=np.float64)
        h = C.c_void_p()
        _lib.check(_lib.lib().cb_mme_create(self.ctx.handle, self.n, self.p, self.q,
                                            _lib.ptr(self._keep[0]), _lib.ptr(self._keep[1]),
                                            _lib.ptr(self._keep[2]), _lib.ptr(aid), _lib.ptr(y),
                                            C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().cb_mme_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def dim(self) -> int:
        return self.p + self.q

    def set_qinv_dense(self, Q: np.ndarray) -> None:
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        _lib.check(_lib.lib().cb_mme_set_qinv_dense(self.handle, _lib.ptr(Q), Q.shape[1]))

    def set_qinv_dense_device(self, dptr: int, ld: int) -> None:
        _lib.check(_lib.lib().cb_mme_set_qinv_dense_device(self.handle, dptr, ld))

    def set_qinv_csr(self, ptr, col, val) -> None:
        ptr = np.ascontiguousarray(ptr, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        _lib.check(_lib.lib().cb_mme_set_qinv_csr(self.handle, _lib.ptr(ptr), _lib.ptr(col), _lib.ptr(val)))

    def set_qinv_split(self, ptr, col, val, gmap, dD: int, ldd: int) -> None:
        """H^-1 = A^-1 (CSR) + scatter(D) on the genotyped rows gmap; D a
        device pointer (ng x ng, leading dimension ldd)."""
        ptr = np.ascontiguousarray(ptr, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        gmap = np.ascontiguousarray(gmap, dtype=np.int32)
        _lib.check(_lib.lib().cb_mme_set_qinv_split(self.handle, _lib.ptr(ptr), _lib.ptr(col), _lib.ptr(val),
                                                    gmap.size, _lib.ptr(gmap), dD, ldd))

    def rhs(self) -> np.ndarray:
        out = np.empty(self.dim)
        _lib.check(_lib.lib().cb_mme_rhs(self.handle, _lib.ptr(out)))
        return out

    def apply(self, lam: float, X: np.ndarray) -> np.ndarray:
        X2 = np.ascontiguousarray(np.asarray(X, dtype=np.float64).reshape(self.dim, -1).T)
        r = X2.shape[0]
        Y = np.empty_like(X2)
        _lib.check(_lib.lib().cb_mme_apply(self.handle, lam, r, _lib.ptr(X2), _lib.ptr(Y)))
        return Y.T.reshape(np.asarray(X).shape)

    def solve(self, lam: float, B: np.ndarray, opts: _lib.SolveOptions | None = None):
        """Solve C(lam) X = B; returns (X, iterations, max relative residual)."""
        opts = opts or solve_options()
        B2 = np.ascontiguousarray(np.asarray(B, dtype=np.float64).reshape(self.dim, -1).T)
        r = B2.shape[0]
        X = np.empty_like(B2)
        it = C.c_int64()
        rel = C.c_double()
        _lib.check(_lib.lib().cb_mme_solve(self.handle, lam, C.byref(opts), r, _lib.ptr(B2), _lib.ptr(X),
                                           C.byref(it), C.byref(rel)))
        return X.T.reshape(np.asarray(B).shape), it.value, rel.value

    def solve_device_noalloc(self, lam: float, dB: int, ldb: int, dX: int, ldx: int, r: int,
                             opts: _lib.SolveOptions | None = None):
        """solve_device with the scratch in the library's fixed device array
        (no device allocation); returns (iterations, max relative residual)."""
        opts = opts or solve_options()
        it = C.c_int64()
        rel = C.c_double()
        _lib.check(_lib.lib().cb_mme_solve_device_noalloc(self.handle, lam, C.byref(opts), r, C.c_void_p(dB), ldb,
                                                          C.c_void_p(dX), ldx, C.byref(it), C.byref(rel)))
        return it.value, rel.value

    def beta_cov(self, lam: float, var_e: float, opts: _lib.SolveOptions | None = None):
        """GSSolve's fixed-effect SEs (mixed_model_solver.cpp:346-417):
        returns (Cinv_beta p x p, beta_se p)."""
        opts = opts or solve_options()
        cinv = np.empty((self.p, self.p))
        se = np.empty(self.p)
        _lib.check(_lib.lib().cb_mme_beta_cov(self.handle, lam, var_e, C.byref(opts), _lib.ptr(cinv), _lib.ptr(se),
                                              None))
        return cinv.T.copy(), se

    def solve_device(self, lam: float, dB: int, ldb: int, dX: int, ldx: int, r: int,
                     opts: _lib.SolveOptions | None = None):
        """solve() on device-resident column blocks (device pointers with
        leading dimensions >= dim, e.g. torch tensors of shape (r, ld));
        returns (iterations, max relative residual)."""
        opts = opts or solve_options()
        it = C.c_int64()
        rel = C.c_double()
        _lib.check(_lib.lib().cb_mme_solve_device(self.handle, lam, C.byref(opts), r, C.c_void_p(dB), ldb,
                                                  C.c_void_p(dX), ldx, C.byref(it), C.byref(rel)))
        return it.value, rel.value


class AIREML:
    """VCESolver for the single-component model (vce.h:223-390)."""

    def __init__(self, model: MixedModel, opts: _lib.RemlOptions | None = None,
                 var_u: float = 1.0, var_e: float = 1.0):
        self.model = model
        self.opts = opts or reml_options()
        self.var_u0, self.var_e0 = var_u, var_e  # phenotype_loader.cpp:174-176 defaults
        self.var_u = self.var_e = None
        self.solution = None
        self.history: list[tuple[float, float]] = []
        self.iterations = 0
        self.converged = False
        self.ai_last = None
        self.var_u_se = self.var_e_se = None

    def setInitialVars(self, var_u: float, var_e: float) -> None:
        self.var_u0, self.var_e0 = var_u, var_e

    def solve(self) -> None:
        m = self.model
        kmax = int(self.opts.max_iter)
        hu = np.zeros(max(1, kmax))
        he = np.zeros(max(1, kmax))
        sol = np.empty(m.dim)
        vu, ve = C.c_double(), C.c_double()
        it = C.c_int64()
        conv = C.c_int()
        ai = np.empty(4)
        se = np.empty(2)
        _lib.check(_lib.lib().cb_reml_run(m.handle, C.byref(self.opts), self.var_u0, self.var_e0,
                                          C.byref(vu), C.byref(ve), C.byref(it), C.byref(conv),
                                          _lib.ptr(hu), _lib.ptr(he), _lib.ptr(sol), _lib.ptr(ai), _lib.ptr(se)))
        self.var_u, self.var_e = vu.value, ve.value
        self.ai_last = ai.reshape(2, 2)
        self.var_u_se, self.var_e_se = float(se[0]), float(se[1])
        self.iterations, self.converged = it.value, bool(conv.value)
        self.history = list(zip(hu[: it.value], he[: it.value]))
        self.solution = sol

    def score(self, var_u: float, var_e: float, want_ai: bool = True):
        m = self.model
        g = np.empty(2)
        ai = np.empty(4)
        tr = C.c_double()
        sol = np.empty(m.dim)
        _lib.check(_lib.lib().cb_reml_score(m.handle, C.byref(self.opts), var_u, var_e, _lib.ptr(g),
                                            _lib.ptr(ai) if want_ai else None, C.byref(tr), _lib.ptr(sol)))
        return g, ai.reshape(2, 2) if want_ai else None, tr.value, sol

    def branch(self) -> dict:
        """The steps, trace rule and logL check this run takes
        (cb_reml_branch; VCESolver::create's size-based choice for
        CB_VCE_AUTO)."""
        f, l, t, g = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lo, hi = C.c_int64(), C.c_int64()
        _lib.check(_lib.lib().cb_reml_branch(self.model.handle, C.byref(self.opts), C.byref(f), C.byref(l),
                                             C.byref(t), C.byref(g), C.byref(lo), C.byref(hi)))
        return {"first": f.value, "later": l.value, "trace": t.value, "logl_checked": bool(g.value),
                "min_probes": lo.value, "max_probes": hi.value}

    def probes(self) -> np.ndarray:
        m = self.model
        b = self.branch()
        k = m.q if b["trace"] == 2 else b["max_probes"] if b["trace"] == 1 else int(self.opts.n_probes)
        out = np.empty((k, m.q))
        _lib.check(_lib.lib().cb_reml_copy_probes(m.handle, C.byref(self.opts), _lib.ptr(out), m.q))
        return out.T

    def logL(self, var_u: float, var_e: float) -> float:
        """calc_logL_internal (vce.cpp:3533-3588) at the given variances."""
        v = C.c_double()
        _lib.check(_lib.lib().cb_reml_logl(self.model.handle, var_u, var_e, C.byref(v), None))
        return v.value

    def calculate_SE(self) -> None:
        """SEs from the last AI matrix; solve() already computed them
        (vce.cpp:3197-3368)."""
        if self.var_u_se is None:
            raise RuntimeError("calculate_SE: run solve() first")

    # VCESolver getters
    def getSEVarsU(self) -> list[float]:
        return [self.var_u_se]

    def getSEVarE(self) -> float:
        return self.var_e_se

    def getVarE(self) -> float:
        return self.var_e

    def getVarsU(self) -> list[float]:
        return [self.var_u]

    def getSolution(self) -> np.ndarray:
        return self.solution
