"""Oracle (test infrastructure) for the MME, PCG and AI-REML part of the hot
path: dense FP64 numpy restatements of the reference algorithms.

  mme_dense            MMELHSBuilder LHS/RHS (src/core/mme_builder.cpp:72-328)
  probes               the probe stream the GPU uses (splitmix64 signs)
  reml_eval            data-driven score + AI (src/core/vce.cpp:2414-2675)
  reml_run             AI_REML::solve loop (vce.cpp:741-837) with the AI step
                       (vce.cpp:2596-2813 semantics for the PCG path) or the EM
                       + Hutchinson step (vce.cpp:1490-1574); variance
                       stabiliser vce.cpp:1370-1389
  logL                 calc_logL_internal (vce.cpp:3533-3588)
  beta_cov             GSSolve's fixed-effect covariance / SEs
                       (mixed_model_solver.cpp:346-417)
  adaptive_trace       AdaptiveTraceEstimator::estimate's stopping rule
                       (adaptive_trace_estimator.cpp:11-83)
  dense_exact_eval     run_dense_exact_step terms (vce.cpp:2859-3120): P from
                       the MME, M_u = P Z A Z', AI = 1/2 [tr(M_u M_u),
                       tr(M_u P); ., ||P||_F^2], exact gradient, logL with
                       log|A|
  dense_exact_run      AI_REML loop with that step and its <= 6 logL halvings
                       (vce.cpp:3121-3195)
  henderson_ainv       invertA_henderson without inbreeding
                       (src/core/a_matrix.cpp:392-439, textbook rules)
  ml_inbreeding        ml() (a_matrix.cpp:127-186): F and d_ii by the dense
                       descending scan, pure Python (small pedigrees)
  hinv_blocks          ssGBLUP split-mode set-up (src/io/relationship_builder.cpp:
                       481-633): A22 from the tabular A, tuneG, blend, G^-1
                       by the eigen-bent inverse, A22^-1, D = tau G^-1 - omega A22^-1
  henderson_ainv_f     invertA_henderson with ml() inbreeding, the
                       reference's triplet order, symmetrisation, diagonal
                       floor and prune (a_matrix.cpp:43-80,392-439)

Pinning (tests/test_oracle_reml.py): MME solution vs the V-form BLUP/GLS
formulas; the score vs central finite differences of logL; EM and AI fixed
points coincide; Henderson A^-1 vs the inverse of the tabular A on small
pedigrees and the 4-animal fixture by hand; ml_inbreeding's F vs
diag(tabular A) - 1 on inbred pedigrees and henderson_ainv_f vs the inverse
of the tabular A there; dense_exact_eval's gradient vs central finite
differences of its logL and its P vs the V-form projection
V^-1 - V^-1 X (X'V^-1 X)^-1 X'V^-1; mv_eigen_vce's fixed point as a
stationary point of mv_reml_logl (central finite differences) and mv_solve
vs the multi-trait V-form BLUP; hinv_blocks' tuneG (b G + a matches A22's
mean diagonal and off-diagonal) and its A22 vs the genotyped block of the
inverse of Henderson's A^-1 (tests/test_pedigree.py); probes vs a scalar
splitmix64.
"""
from __future__ import annotations

import numpy as np

GOLD = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + GOLD
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def probes(q: int, k: int, seed: int, exact: bool = False) -> np.ndarray:
    """q x k probe matrix: sign of splitmix64(seed + g*(c*q + j + 1))."""
    if exact:
        return np.eye(q)
    c = np.arange(k, dtype=np.uint64)[None, :]
    j = np.arange(q, dtype=np.uint64)[:, None]
    with np.errstate(over="ignore"):
        arg = np.uint64(seed) + GOLD * (c * np.uint64(q) + j + np.uint64(1))
    h = splitmix64(arg)
    return np.where((h >> np.uint64(63)) != 0, -1.0, 1.0)


def design_w(X: np.ndarray, aid: np.ndarray, q: int) -> np.ndarray:
    n = X.shape[0]
    Z = np.zeros((n, q))
    ok = aid >= 0
    Z[np.nonzero(ok)[0], aid[ok]] = 1.0
    return np.hstack([X, Z])


def mme_dense(X, aid, y, Q, lam):
    """C = W'W + lam blockdiag(0, Q), rhs = W'y (data scale, as build_lhs).
    W'W is formed from a sparse copy of W (one level per record), so the
    product costs O(nnz) rather than a dense n x (p+q) GEMM."""
    import scipy.sparse as sp

    p, q = X.shape[1], Q.shape[0]
    W = design_w(X, aid, q)
    Ws = sp.csr_matrix(W)
    C = (Ws.T @ Ws).toarray()
    C[p:, p:] += lam * Q
    return C, Ws.T @ y, W


def adaptive_trace(t, min_k, max_k):
    """AdaptiveTraceEstimator::estimate (src/core/adaptive_trace_estimator.cpp:
    11-83) over per-probe values t: probes in batches of 16 (one thread);
    after each batch, once at least min_k are done, stop when SE/|mean| <
    0.05; at most max_k. Returns (mean, probes used)."""
    done = 0
    while done < max_k:
        done = min(max_k, done + 16)
        if done >= min_k and done >= 2:
            v = np.asarray(t[:done], dtype=np.float64)
            ssum, sq = float(np.sum(v)), float(np.sum(v * v))
            mean = ssum / done
            var = (sq - ssum * ssum / done) / (done - 1.0)
            se = np.sqrt(max(0.0, var) / done)
            if abs(mean) > 1e-12 and se / abs(mean) < 0.05:
                break
    return float(np.mean(np.asarray(t[:done], dtype=np.float64))), done


def reml_eval(X, aid, y, Q, vu, ve, Zp, want_ai=True, exact=False, adaptive=None):
    """Score, AI and trace at (vu, ve) with probe matrix Zp (q x k);
    adaptive = (min_k, max_k) applies adaptive_trace to Zp's columns."""
    n, p = X.shape
    q = Q.shape[0]
    lam = ve / max(vu, 1e-9)
    C, rhs, W = mme_dense(X, aid, y, Q, lam)
    k = Zp.shape[1]
    B = np.zeros((p + q, 1 + k))
    B[:, 0] = rhs
    B[p:, 1:] = Zp
    S = np.linalg.solve(C, B)
    u = S[p:, 0]
    e = y - W @ S[:, 0]
    esse = float(e @ e)
    uQu = float(u @ Q @ u)
    QZ = Q @ Zp
    t = np.einsum("ij,ij->j", QZ, S[p:, 1:])
    if exact:
        tr = float(t.sum())
    elif adaptive is not None:
        tr, _ = adaptive_trace(t, adaptive[0], adaptive[1])
    else:
        tr = float(t.mean())
    lam_tr = min(lam * tr, q - 1e-6)
    grad = np.array([0.5 * (uQu + tr * ve - q * vu) / (vu * vu),
                     0.5 * (esse - ve * (n - p - q + lam_tr)) / (ve * ve)])
    out = {"grad": grad, "tr": tr, "uQu": uQu, "esse": esse, "sol": S[:, 0]}
    if want_ai:
        v_u = np.zeros(n)
        ok = aid >= 0
        v_u[ok] = u[aid[ok]] / vu
        v_e = e / ve
        S2 = np.linalg.solve(C, W.T @ np.stack([v_u, v_e], axis=1))
        T = W @ S2
        Pv_u = (v_u - T[:, 0]) / ve
        Pv_e = (v_e - T[:, 1]) / ve
        ai = 0.5 * np.array([[v_u @ Pv_u, v_u @ Pv_e], [v_u @ Pv_e, v_e @ Pv_e]])
        out["ai"] = ai
    return out


def project_spd_2x2(A, eps_rel=1e-8, eps_abs=1e-12):
    """vce.cpp:267-283."""
    S = 0.5 * (A + A.T)
    w, V = np.linalg.eigh(S)
    floor = max(eps_abs, eps_rel * max(w.max(), 0.0))
    w = np.maximum(w, floor)
    return (V * w) @ V.T


def reml_run(X, aid, y, Q, vu0=1.0, ve0=1.0, algorithm="ai", max_iter=20, tol=1e-4, Zp=None,
             exact=False, later=None, adaptive=None, logl_check=False):
    """AI_REML::solve (vce.cpp:741-837): `algorithm` is the step of the
    first iteration ("ai" or "em"), `later` (default: the same) that of the
    iterations after it; adaptive = (min_k, max_k) traces (reml_eval); with
    logl_check an AI step is accepted only where logL improves (<= 15
    halvings, positive, within 5x, vce.cpp:2797-2848), unstabilised, and a
    failed one is an EM step."""
    n, p = X.shape
    q = Q.shape[0]
    ymean = y.mean()
    yvar = max(1e-12, float(((y - ymean) ** 2).sum() / (n - 1))) if n > 1 else max(1e-12, float(y @ y))
    vfloor, vceil = 1e-12, max(1.0, yvar * 1e6)

    def stabilize(prop, prev):
        if not np.isfinite(prop) or prop <= 0:
            return max(prev, vfloor)
        lo, hi = vfloor, vceil
        if np.isfinite(prev) and prev > 0:
            lo, hi = max(lo, prev * 0.1), min(hi, prev * 10.0)
        if lo > hi:
            lo = vfloor
        return min(max(prop, lo), hi)

    vu, ve = vu0, ve0
    hist = []
    conv = False
    ai_last = None
    later = later or algorithm
    for it in range(max_iter):
        ou, oe = vu, ve
        step = algorithm if it == 0 else later
        s = reml_eval(X, aid, y, Q, vu, ve, Zp, want_ai=(step == "ai"), exact=exact, adaptive=adaptive)
        lam = ve / max(vu, 1e-9)
        em_u = (s["uQu"] + s["tr"] * ve) / q
        em_e = (s["esse"] + ve * ((p + q) - lam * s["tr"])) / n
        if step == "em":
            vu, ve = stabilize(em_u, ou), stabilize(em_e, oe)
        else:
            A = project_spd_2x2(s["ai"])
            ai_last = A
            d = np.linalg.solve(A, s["grad"])
            ok = np.all(np.isfinite(d)) and abs(d[0]) <= 100 * vu and abs(d[1]) <= 100 * ve
            if ok and logl_check:
                cur = logL(X, aid, y, Q, vu, ve)
                ok = False
                for h in range(15):
                    nu, ne = vu + d[0], ve + d[1]
                    if nu > 0 and ne > 0 and abs(d[0]) <= 5 * vu and abs(d[1]) <= 5 * ve:
                        nl = logL(X, aid, y, Q, nu, ne)
                        if np.isfinite(nl) and nl > cur:
                            ok = True
                            break
                    d = d * 0.5
                if ok:
                    vu, ve = nu, ne
                else:
                    vu, ve = stabilize(em_u, ou), stabilize(em_e, oe)
            else:
                if ok:
                    for h in range(15):
                        if vu + d[0] > 0 and ve + d[1] > 0 and abs(d[0]) <= 5 * vu and abs(d[1]) <= 5 * ve:
                            break
                        d = d * 0.5
                    else:
                        ok = False
                if ok:
                    vu, ve = stabilize(vu + d[0], ou), stabilize(ve + d[1], oe)
                else:
                    vu, ve = stabilize(em_u, ou), stabilize(em_e, oe)
        hist.append((vu, ve))
        if abs(ve - oe) + abs(vu - ou) < tol:
            conv = True
            break
    if ai_last is None:  # EM run: the AI matrix at the estimates
        ai_last = project_spd_2x2(reml_eval(X, aid, y, Q, vu, ve, Zp, want_ai=True, exact=exact,
                                            adaptive=adaptive)["ai"])
    # calculate_SE (vce.cpp:3342-3352): SE = sqrt(max(0, diag(AI^-1)))
    se = np.sqrt(np.maximum(0.0, np.diag(np.linalg.inv(ai_last))))
    lam = ve / max(vu, 1e-9)
    C, rhs, _ = mme_dense(X, aid, y, Q, lam)
    sol = np.linalg.solve(C, rhs)
    return {"vu": vu, "ve": ve, "hist": hist, "converged": conv, "sol": sol, "ai_last": ai_last, "se": se}


def beta_cov(X, aid, y, Q, vu, ve):
    """GSSolve's calc_se (src/solve/mixed_model_solver.cpp:346-417): column i
    of C^-1 for each fixed effect, Cinv_beta = its first p rows, beta_se_i =
    sqrt(max(0, ve (C^-1)_ii)), C = C(ve / vu)."""
    p = X.shape[1]
    C, _, _ = mme_dense(X, aid, y, Q, ve / vu)
    E = np.zeros((C.shape[0], p))
    E[np.arange(p), np.arange(p)] = 1.0
    Ci = np.linalg.solve(C, E)[:p]
    return Ci, np.sqrt(np.maximum(0.0, ve * np.diag(Ci)))


def logL(X, aid, y, Q, vu, ve):
    """-1/2 (log|C| + (n-p-q) log ve + q log vu + y'Py), y'Py = (y'y - sol'rhs)/ve."""
    n, p = X.shape
    q = Q.shape[0]
    lam = ve / vu
    C, rhs, _ = mme_dense(X, aid, y, Q, lam)
    sign, logdet = np.linalg.slogdet(C)
    sol = np.linalg.solve(C, rhs)
    yPy = (y @ y - sol @ rhs) / ve
    return -0.5 * (logdet + (n - p - q) * np.log(ve) + q * np.log(vu) + yPy)


def henderson_ainv(sire: np.ndarray, dam: np.ndarray) -> np.ndarray:
    """Dense A^-1 by Henderson's rules without inbreeding (0-based parents,
    -1 unknown); parents must precede offspring."""
    nanim = sire.size
    A = np.zeros((nanim, nanim))
    for i in range(nanim):
        s, d = sire[i], dam[i]
        par = [x for x in (s, d) if x >= 0]
        if len(par) == 2:
            alpha = 2.0
        elif len(par) == 1:
            alpha = 4.0 / 3.0
        else:
            alpha = 1.0
        idx = [i] + par
        coef = [1.0] + [-0.5] * len(par)
        for a, ca in zip(idx, coef):
            for b, cb in zip(idx, coef):
                A[a, b] += alpha * ca * cb
    return A


def tabular_a(sire: np.ndarray, dam: np.ndarray) -> np.ndarray:
    """Numerator relationship matrix by the tabular method."""
    nanim = sire.size
    A = np.zeros((nanim, nanim))
    for i in range(nanim):
        s, d = sire[i], dam[i]
        for j in range(i):
            v = 0.0
            if s >= 0:
                v += 0.5 * A[j, s]
            if d >= 0:
                v += 0.5 * A[j, d]
            A[i, j] = A[j, i] = v
        A[i, i] = 1.0 + (0.5 * A[s, d] if s >= 0 and d >= 0 else 0.0)
    return A


def ml_inbreeding(dam, sire):
    """F and d_ii as ml() computes them (a_matrix.cpp:127-186): for an
    animal with both parents known, L starts at 1/2 on each parent and the
    dense scan j = max(s, d) .. 0 accumulates L_j^2 d_j while passing half of
    L_j to each parent of j; F_i = sum + d_i - 1, clipped at 0."""
    n = len(dam)
    F = [0.0] * n
    dii = [1.0] * n
    L = [0.0] * n
    for i in range(n):
        s, d = int(sire[i]), int(dam[i])
        s_ok, d_ok = 0 <= s < n, 0 <= d < n
        Fs = F[s] if s_ok else 0.0
        Fd = F[d] if d_ok else 0.0
        if s_ok and d_ok:
            dii[i] = 1.0 - 0.25 * (1.0 + Fs) - 0.25 * (1.0 + Fd)
        elif s_ok:
            dii[i] = 1.0 - 0.25 * (1.0 + Fs)
        elif d_ok:
            dii[i] = 1.0 - 0.25 * (1.0 + Fd)
        else:
            dii[i] = 1.0
        if s_ok and d_ok:
            L[s] += 0.5
            L[d] += 0.5
            acc = 0.0
            for j in range(max(s, d), -1, -1):
                if L[j] > 0.0:
                    acc += L[j] * L[j] * dii[j]
                    sj, dj = int(sire[j]), int(dam[j])
                    if 0 <= sj < n:
                        L[sj] += 0.5 * L[j]
                    if 0 <= dj < n:
                        L[dj] += 0.5 * L[j]
                    L[j] = 0.0
            f = acc + dii[i] - 1.0
            F[i] = f if f >= 0.0 else 0.0
        else:
            F[i] = 0.0
    return np.array(F), np.array(dii)


def henderson_ainv_f(dam, sire):
    """Dense A^-1 as invertA_henderson builds it (a_matrix.cpp:392-439):
    parents outside [0, n) become unknown, F and d_ii from ml_inbreeding,
    triplets summed in insertion order, 0.5 (M + M^T), diagonal < 1e-15 ->
    1e-12, entries with |v| <= 1e-24 dropped (zero)."""
    n = len(dam)
    d = [int(x) if 0 <= int(x) < n else -1 for x in dam]
    s = [int(x) if 0 <= int(x) < n else -1 for x in sire]
    F, dii = ml_inbreeding(d, s)
    M = np.zeros((n, n))
    for i in range(n):
        di = 1.0 / dii[i]
        trip = [(i, i, di)]
        if s[i] >= 0:
            trip += [(i, s[i], -0.5 * di), (s[i], i, -0.5 * di), (s[i], s[i], 0.25 * di)]
        if d[i] >= 0:
            trip += [(i, d[i], -0.5 * di), (d[i], i, -0.5 * di), (d[i], d[i], 0.25 * di)]
        if s[i] >= 0 and d[i] >= 0:
            trip += [(s[i], d[i], 0.25 * di), (d[i], s[i], 0.25 * di)]
        for r, c, v in trip:  # row-major accumulation in insertion order per entry
            M[r, c] += v
    M = 0.5 * (M + M.T)
    dg = np.diag(M).copy()
    dg[dg < 1e-15] = 1e-12
    np.fill_diagonal(M, dg)
    M[np.abs(M) <= 1e-24] = 0.0
    return M, F, dii


def hinv_blocks(dam, sire, gmap, G, tune_g=True, alpha=0.05, tau=1.0, omega=1.0, ridge=0.0):
    """relationship_builder.cpp:481-633 on a dense A (tabular method; small
    pedigrees). Returns (D, A22, (a, b))."""
    from . import grm as og

    A = tabular_a(np.asarray(sire), np.asarray(dam))
    ng = len(gmap)
    A22 = np.eye(ng)
    for i in range(ng):
        for j in range(ng):
            if gmap[i] >= 0 and gmap[j] >= 0:
                A22[i, j] = A[gmap[i], gmap[j]]
            else:
                A22[i, j] = 1.0 if i == j else 0.0
    G = np.array(G, dtype=np.float64, copy=True)
    a, b = 0.0, 1.0
    if tune_g:
        off = ~np.eye(ng, dtype=bool)
        mgd, mad = np.trace(G) / ng, np.trace(A22) / ng
        noff = ng * (ng - 1.0)
        mgo = G[off].sum() / noff if noff > 0 else 0.0
        mao = A22[off].sum() / noff if noff > 0 else 0.0
        if abs(mgd - mgo) > 1e-12:
            b = (mad - mao) / (mgd - mgo)
        a = mad - b * mgd
        G = b * G + a
    G = (1.0 - alpha) * G + alpha * A22
    Ginv = og.invert_psd_matrix(G, ridge)
    A22r = A22 + ridge * np.eye(ng) if ridge > 0 else A22
    A22inv = np.linalg.inv(A22r)
    return tau * Ginv - omega * A22inv, A22, (a, b)


def dense_exact_eval(X, aid, y, Q, vu, ve):
    """The terms of run_dense_exact_step (vce.cpp:2859-3120) at (vu, ve)."""
    n, p = X.shape
    q = Q.shape[0]
    A = np.linalg.inv(Q)
    sign, ldQ = np.linalg.slogdet(Q)
    log_det_A = -ldQ
    lam = ve / vu
    C, rhs, W = mme_dense(X, aid, y, Q, lam)
    sign, ldC = np.linalg.slogdet(C)
    sol = np.linalg.solve(C, rhs)
    e = y - W @ sol
    logL = -0.5 * (ldC + (n - p - q) * np.log(ve) + q * np.log(vu) + y @ e / ve + log_det_A)
    H = W @ np.linalg.solve(C, W.T)
    P = (np.eye(n) - H) / ve
    Z = W[:, p:]
    ZAZ = Z @ A @ Z.T
    Mu = P @ ZAZ
    ai = 0.5 * np.array([[np.sum(Mu * Mu.T), np.sum(Mu * P)], [np.sum(Mu * P), np.sum(P * P)]])
    Py = e / ve
    grad = np.array([-0.5 * np.trace(Mu) + 0.5 * Py @ ZAZ @ Py, -0.5 * np.trace(P) + 0.5 * Py @ Py])
    return {"logL": logL, "grad": grad, "ai": ai, "P": P, "sol": sol}


def dense_exact_run(X, aid, y, Q, vu0=1.0, ve0=1.0, max_iter=20, tol=1e-4):
    """AI_REML::solve with run_dense_exact_step (vce.cpp:741-837,2949-3195):
    Newton on the projected AI, up to 6 halvings, accepting a positive step
    that raises logL or, at the last halving, any positive step. A failed
    step stops the run (unconverged)."""
    vu, ve = vu0, ve0
    hist = []
    conv = False
    ai_last = None
    for _ in range(max_iter):
        t = dense_exact_eval(X, aid, y, Q, vu, ve)
        A = project_spd_2x2(t["ai"])
        ai_last = A
        delta = np.linalg.solve(A, t["grad"])
        ok = False
        for k in range(6):
            nu, ne = vu + delta[0], ve + delta[1]
            if nu > 0 and ne > 0:
                nl = dense_exact_eval(X, aid, y, Q, nu, ne)["logL"]
                if nl > t["logL"] or k == 5:
                    ok = True
                    break
            delta = delta * 0.5
        if not ok:
            break
        diff = abs(ne - ve) + abs(nu - vu)
        vu, ve = nu, ne
        hist.append((vu, ve))
        if diff < tol:
            conv = True
            break
    se = np.sqrt(np.maximum(0.0, np.diag(np.linalg.inv(ai_last))))
    lam = ve / vu
    C, rhs, _ = mme_dense(X, aid, y, Q, lam)
    return {"vu": vu, "ve": ve, "hist": hist, "converged": conv, "sol": np.linalg.solve(C, rhs),
            "ai_last": ai_last, "se": se}


def mv_eigen_vce(Y, X, G, max_iter=500, tol=1e-4):
    """MvEigenVce with covariate projection (src/core/eigen_vce.cpp:106-262,
    options of variance_estimator.cpp:289-294): A = null space of X' from a
    QR of X, eigen-rotation of A'GA, EM on the rotated rows; start
    Vg = Ve = Cov(Y)/2 with the divisor n - c - 1 the reference uses; stop
    when both relative Frobenius changes are below tol. Returns (Vg, Ve,
    iterations, converged)."""
    n, t = Y.shape
    c = X.shape[1]
    Qf, _ = np.linalg.qr(X, mode="complete")
    A = Qf[:, c:]
    Yr = A.T @ Y
    Gr = A.T @ G @ A
    D, U = np.linalg.eigh(Gr)
    D = np.where(D < 0, 0.0, D)
    Uty = U.T @ Yr
    ns = n - c
    Yc = Y - Y.mean(axis=0)
    Cov = Yc.T @ Yc / (ns - 1)
    Vg, Ve = 0.5 * Cov, 0.5 * Cov
    it = 0
    conv = False
    for it in range(1, max_iter + 1):
        # the per-row updates of updateEM, batched over the rotated rows
        Vi = D[:, None, None] * Vg[None] + Ve[None]
        Pi = np.linalg.inv(Vi)
        Wg = D[:, None, None] * np.einsum("ab,ibc->iac", Vg, Pi)
        uh = np.einsum("iab,ib->ia", Wg, Uty)
        ok = D > 1e-12
        Sg = (np.einsum("ia,ib->ab", uh[ok], uh[ok] / D[ok, None]) + ok.sum() * Vg
              - np.einsum("iab,bc->ac", Wg[ok], Vg) + (~ok).sum() * Vg)
        We = np.einsum("ab,ibc->iac", Ve, Pi)
        eh = np.einsum("iab,ib->ia", We, Uty)
        Se = np.einsum("ia,ib->ab", eh, eh) + ns * Ve - np.einsum("iab,bc->ac", We, Ve)
        Vg_new, Ve_new = Sg / ns, Se / ns
        dg = np.linalg.norm(Vg_new - Vg) / max(np.linalg.norm(Vg), 1e-12)
        de = np.linalg.norm(Ve_new - Ve) / max(np.linalg.norm(Ve), 1e-12)
        Vg, Ve = Vg_new, Ve_new
        if dg < tol and de < tol:
            conv = True
            break
    return Vg, Ve, it, conv


def mv_solve(Y, X, Ginv, Vg, Ve):
    """GSSolve::solveMultiTrait (src/solve/mixed_model_solver.cpp:425-524):
    the dense Kronecker MME with unknowns trait-major; returns (beta p x t,
    u n x t)."""
    n, t = Y.shape
    p = X.shape[1]
    Vei, Vgi = np.linalg.inv(Ve), np.linalg.inv(Vg)
    C11 = np.kron(Vei, X.T @ X)
    C12 = np.kron(Vei, X.T)
    C22 = np.kron(Vei, np.eye(n)) + np.kron(Vgi, Ginv)
    L = np.block([[C11, C12], [C12.T, C22]])
    rhs = np.concatenate([(X.T @ Y @ Vei.T).T.ravel(), (Y @ Vei.T).T.ravel()])
    s = np.linalg.solve(L, rhs)
    return s[:p * t].reshape(t, p).T, s[p * t:].reshape(t, n).T


def mv_reml_logl(Y, X, G, Vg, Ve):
    """Multivariate REML log-likelihood (up to a constant) of vec(Y) with
    V = Vg (x) G + Ve (x) I: -1/2 (log|V| + log|X~' V^-1 X~| + y'Py),
    X~ = I_t (x) X, unknowns trait-major."""
    n, t = Y.shape
    V = np.kron(Vg, G) + np.kron(Ve, np.eye(n))
    Xt = np.kron(np.eye(t), X)
    y = Y.T.ravel()
    Vi = np.linalg.inv(V)
    XVX = Xt.T @ Vi @ Xt
    P = Vi - Vi @ Xt @ np.linalg.solve(XVX, Xt.T @ Vi)
    return -0.5 * (np.linalg.slogdet(V)[1] + np.linalg.slogdet(XVX)[1] + y @ P @ y)
