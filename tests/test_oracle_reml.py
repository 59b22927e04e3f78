"""Pins the MME / REML oracle (oracle/reml.py) against facts that do not
restate it: V-form BLUP/GLS, central finite differences of the REML
log-likelihood, the EM = AI fixed point, Henderson's A^-1 against the inverse
of the tabular A, the reference's 4-animal pedigree fixture, and the probe
stream against a scalar splitmix64."""
import numpy as np
import pytest

from cosmicblup_amd import design, synth
from oracle import grm as og
from oracle import reml as orl


def small_problem(n=60, m=300, ngroups=4, seed=3):
    x, _ = synth.simulate_dosages(n, m, maf=(0.05, 0.5), seed=seed)
    bed = synth.encode_bed(x)
    G, _, _ = og.grm_vanraden(bed, n, m)
    Q = og.invert_psd_matrix(G)
    y, cg = synth.simulate_phenotype(x, n_qtl=30, h2=0.4, n_groups=ngroups, seed=seed + 1)
    fd = design.build_fixed_design([[c] for c in cg])
    X = design.dense_x(fd)
    aid = np.arange(n, dtype=np.int32)
    return X, aid, y, Q, G


def test_design_levels_string_sorted_first_dropped():
    fd = design.build_fixed_design([["b"], ["a"], ["c"], ["a"], ["10"], ["9"]])
    # std::set<string> order: "10" < "9" < "a" < "b" < "c"; "10" is the base
    assert fd.names == ["mu", "factor1_9", "factor1_a", "factor1_b", "factor1_c"]
    X = design.dense_x(fd)
    assert X[4].tolist() == [1, 0, 0, 0, 0]
    assert X[0].tolist() == [1, 0, 0, 1, 0]


def test_mme_solution_equals_vform_blup():
    X, aid, y, Q, G = small_problem()
    vu, ve = 0.7, 1.3
    C, rhs, W = orl.mme_dense(X, aid, y, Q, ve / vu)
    sol = np.linalg.lstsq(C, rhs, rcond=None)[0]
    Gb = og.bent_matrix(G)  # Q^-1
    V = vu * Gb + ve * np.eye(len(y))
    Vi = np.linalg.inv(V)
    XtViX = X.T @ Vi @ X
    b = np.linalg.lstsq(XtViX, X.T @ Vi @ y, rcond=None)[0]
    u = vu * Gb @ Vi @ (y - X @ b)
    np.testing.assert_allclose(sol[X.shape[1]:], u, atol=1e-7 * np.abs(u).max())
    np.testing.assert_allclose(X @ sol[:X.shape[1]], X @ b, atol=1e-7)


def test_score_matches_central_fd_of_logL():
    X, aid, y, Q, _ = small_problem()
    vu, ve = 0.6, 0.9
    s = orl.reml_eval(X, aid, y, Q, vu, ve, np.eye(Q.shape[0]), exact=True)
    h = 1e-4
    fd_u = (orl.logL(X, aid, y, Q, vu * (1 + h), ve) - orl.logL(X, aid, y, Q, vu * (1 - h), ve)) / (2 * h * vu)
    fd_e = (orl.logL(X, aid, y, Q, vu, ve * (1 + h)) - orl.logL(X, aid, y, Q, vu, ve * (1 - h))) / (2 * h * ve)
    np.testing.assert_allclose(s["grad"], [fd_u, fd_e], rtol=1e-5, atol=1e-6)


def test_ai_matrix_is_spd_average_information():
    # AI_ij = 1/2 v_i' P v_j with P positive semidefinite -> SPD 2x2
    X, aid, y, Q, _ = small_problem()
    s = orl.reml_eval(X, aid, y, Q, 0.6, 0.9, np.eye(Q.shape[0]), exact=True)
    assert np.all(np.linalg.eigvalsh(s["ai"]) > 0)
    assert np.allclose(s["ai"], s["ai"].T)


def test_em_and_ai_share_the_fixed_point():
    X, aid, y, Q, _ = small_problem()
    Zp = orl.probes(Q.shape[0], 16, seed=5)
    ai = orl.reml_run(X, aid, y, Q, algorithm="ai", max_iter=50, tol=1e-10, Zp=Zp)
    em = orl.reml_run(X, aid, y, Q, algorithm="em", max_iter=3000, tol=1e-12, Zp=Zp)
    assert ai["converged"]
    np.testing.assert_allclose([ai["vu"], ai["ve"]], [em["vu"], em["ve"]], rtol=2e-5)
    s = orl.reml_eval(X, aid, y, Q, ai["vu"], ai["ve"], Zp, want_ai=False)
    assert np.abs(s["grad"]).max() < 1e-5


def test_probe_stream_scalar():
    def sm(x):
        M = (1 << 64) - 1
        x = (x + 0x9E3779B97F4A7C15) & M
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M
        return x ^ (x >> 31)

    q, k, seed = 7, 3, 42
    Z = orl.probes(q, k, seed)
    for c in range(k):
        for j in range(q):
            h = sm((seed + 0x9E3779B97F4A7C15 * (c * q + j + 1)) & ((1 << 64) - 1))
            assert Z[j, c] == (-1.0 if h >> 63 else 1.0)
    assert set(np.unique(orl.probes(50, 8, 1))) == {-1.0, 1.0}


def test_hutchinson_unbiased():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((40, 40))
    A = A @ A.T
    Zp = orl.probes(40, 4000, seed=9)
    est = np.einsum("ij,ij->j", Zp, A @ Zp).mean()
    assert abs(est - np.trace(A)) < 4 * np.sqrt(2 * (A * A).sum() / 4000)


def test_henderson_ainv_inverts_tabular_a():
    rng = np.random.default_rng(4)
    nan = 40
    sire = np.full(nan, -1)
    dam = np.full(nan, -1)
    for i in range(10, nan):
        sire[i] = rng.integers(0, 10) if rng.random() > 0.1 else -1
        dam[i] = rng.integers(0, 10) if rng.random() > 0.1 else -1
        if sire[i] == dam[i] and sire[i] >= 0:
            dam[i] = -1
    A = orl.tabular_a(sire, dam)
    # no inbreeding here (parents are unrelated founders)
    np.testing.assert_allclose(orl.henderson_ainv(sire, dam) @ A, np.eye(nan), atol=1e-12)


def test_pblup_minimal_fixture_ainv_by_hand():
    # tests/fixtures/pblup_minimal/pedigree.txt: id3, id4 full sibs of id1 x id2
    sire = np.array([-1, -1, 1, 1])   # column "sire" = id2
    dam = np.array([-1, -1, 0, 0])    # column "dam" = id1
    Ai = orl.henderson_ainv(sire, dam)
    expect = np.array([[2, 1, -1, -1], [1, 2, -1, -1], [-1, -1, 2, 0], [-1, -1, 0, 2]], dtype=float)
    np.testing.assert_allclose(Ai, expect)


def _small_gblup(n=120, m=500, seed=3):
    from cosmicblup_amd import design, synth
    from oracle import grm as og

    x, _ = synth.simulate_dosages(n, m, maf=(0.05, 0.5), seed=seed)
    G, _, _ = og.grm_vanraden(synth.encode_bed(x), n, m)
    Q = og.invert_psd_matrix(G + 0.05 * np.eye(n))
    y, cg = synth.simulate_phenotype(x, n_qtl=30, h2=0.4, n_groups=4, seed=seed + 1)
    fd = design.build_fixed_design([[c] for c in cg])
    return design.dense_x(fd), np.arange(n, dtype=np.int32), y, Q


def test_dense_exact_gradient_is_fd_of_logL():
    X, aid, y, Q = _small_gblup()
    vu, ve = 0.6, 0.9
    t = orl.dense_exact_eval(X, aid, y, Q, vu, ve)
    h = 1e-5
    fu = (orl.dense_exact_eval(X, aid, y, Q, vu * (1 + h), ve)["logL"]
          - orl.dense_exact_eval(X, aid, y, Q, vu * (1 - h), ve)["logL"]) / (2 * h * vu)
    fe = (orl.dense_exact_eval(X, aid, y, Q, vu, ve * (1 + h))["logL"]
          - orl.dense_exact_eval(X, aid, y, Q, vu, ve * (1 - h))["logL"]) / (2 * h * ve)
    np.testing.assert_allclose(t["grad"], [fu, fe], rtol=1e-6)


def test_dense_exact_P_is_the_V_form_projection():
    X, aid, y, Q = _small_gblup(n=80, m=300)
    vu, ve = 0.7, 1.1
    P = orl.dense_exact_eval(X, aid, y, Q, vu, ve)["P"]
    V = vu * np.linalg.inv(Q) + ve * np.eye(y.size)
    Vi = np.linalg.inv(V)
    PV = Vi - Vi @ X @ np.linalg.solve(X.T @ Vi @ X, X.T @ Vi)
    np.testing.assert_allclose(P, PV, atol=1e-10)


def test_dense_exact_run_reaches_the_score_zero():
    X, aid, y, Q = _small_gblup()
    r = orl.dense_exact_run(X, aid, y, Q, max_iter=50, tol=1e-8)
    assert r["converged"]
    g = orl.dense_exact_eval(X, aid, y, Q, r["vu"], r["ve"])["grad"]
    assert np.abs(g).max() < 1e-4 * np.abs(orl.dense_exact_eval(X, aid, y, Q, 1.0, 1.0)["grad"]).max()


def _mv_problem(n=60, t=2, seed=4):
    from cosmicblup_amd import synth
    from oracle import grm as og

    rng = np.random.default_rng(seed)
    x, _ = synth.simulate_dosages(n, 400, maf=(0.1, 0.5), seed=seed)
    G, _, _ = og.grm_vanraden(synth.encode_bed(x), n, 400)
    G = G + 0.05 * np.eye(n)
    Lg = np.linalg.cholesky(G)
    Vg = np.array([[1.0, 0.4], [0.4, 0.8]])[:t, :t]
    Ve = np.array([[1.0, 0.2], [0.2, 1.2]])[:t, :t]
    U = Lg @ rng.standard_normal((n, t)) @ np.linalg.cholesky(Vg).T
    E = rng.standard_normal((n, t)) @ np.linalg.cholesky(Ve).T
    X = np.column_stack([np.ones(n), rng.integers(0, 2, n)])
    Y = X @ rng.standard_normal((2, t)) + U + E
    return Y, X, G


def test_mv_em_fixed_point_is_a_reml_stationary_point():
    """At the EM fixed point every partial derivative of the multivariate REML
    logL vanishes (central differences on each (co)variance parameter)."""
    Y, X, G = _mv_problem()
    Vg, Ve, it, conv = orl.mv_eigen_vce(Y, X, G, max_iter=200000, tol=1e-12)
    assert conv
    h = 1e-6
    for M in ("g", "e"):
        for (a, b) in ((0, 0), (0, 1), (1, 1)):
            E = np.zeros((2, 2))
            E[a, b] = E[b, a] = h
            up = orl.mv_reml_logl(Y, X, G, Vg + (E if M == "g" else 0), Ve + (E if M == "e" else 0))
            dn = orl.mv_reml_logl(Y, X, G, Vg - (E if M == "g" else 0), Ve - (E if M == "e" else 0))
            assert abs((up - dn) / (2 * h)) < 1e-5


def test_mv_solve_is_the_v_form_blup():
    Y, X, G = _mv_problem(n=40)
    Vg = np.array([[0.9, 0.3], [0.3, 0.7]])
    Ve = np.array([[1.1, 0.1], [0.1, 0.9]])
    beta, u = orl.mv_solve(Y, X, np.linalg.inv(G), Vg, Ve)
    n, t = Y.shape
    V = np.kron(Vg, G) + np.kron(Ve, np.eye(n))
    Xt = np.kron(np.eye(t), X)
    y = Y.T.ravel()
    Vi = np.linalg.inv(V)
    b = np.linalg.solve(Xt.T @ Vi @ Xt, Xt.T @ Vi @ y)
    uv = np.kron(Vg, G) @ Vi @ (y - Xt @ b)
    np.testing.assert_allclose(beta.T.ravel(), b, atol=1e-9)
    np.testing.assert_allclose(u.T.ravel(), uv, atol=1e-9)


def test_logl_equals_v_form_reml_plus_log_det_a():
    """calc_logL_internal's MME form against the V-form REML likelihood
    -1/2 (log|V| + log|X'V^-1 X| + y'Py), V = vu Z A Z' + ve I, A = Q^-1:
    the two differ by exactly 1/2 log|A| (the reference omits that
    constant), at several variance pairs."""
    X, aid, y, Q, _ = small_problem(n=90, seed=12)
    # two records share a level and one record has none
    aid = aid.copy()
    aid[5], aid[6] = aid[4], -1
    A = np.linalg.inv(Q)
    Z = orl.design_w(np.zeros((X.shape[0], 0)), aid, Q.shape[0])
    half_logdet_a = 0.5 * np.linalg.slogdet(A)[1]
    for vu, ve in [(0.7, 1.3), (2.0, 0.4), (0.05, 3.0)]:
        V = vu * Z @ A @ Z.T + ve * np.eye(X.shape[0])
        Vi = np.linalg.inv(V)
        XVX = X.T @ Vi @ X
        P = Vi - Vi @ X @ np.linalg.solve(XVX, X.T @ Vi)
        v_form = -0.5 * (np.linalg.slogdet(V)[1] + np.linalg.slogdet(XVX)[1] + y @ P @ y)
        assert abs(orl.logL(X, aid, y, Q, vu, ve) - v_form - half_logdet_a) < 1e-8


def test_adaptive_trace_rule_by_brute_force():
    """adaptive_trace against the stopping rule evaluated independently with
    numpy's two-pass standard deviation at every batch end (16, 32, ...,
    max_k): the first batch end >= min_k with std/sqrt(n)/|mean| < 0.05."""
    rng = np.random.default_rng(5)
    for loc, scale, lo, hi in [(10.0, 1.0, 30, 100), (1.0, 3.0, 33, 100), (1.0, 20.0, 30, 100),
                               (-5.0, 8.0, 30, 150), (2.0, 0.0, 30, 100)]:
        t = loc + scale * rng.standard_normal(hi)
        want = hi
        for end in list(range(16, hi, 16)) + [hi]:
            if end < lo:
                continue
            v = t[:end]
            if abs(v.mean()) > 1e-12 and np.std(v, ddof=1) / np.sqrt(end) / abs(v.mean()) < 0.05:
                want = end
                break
        mean, used = orl.adaptive_trace(t, lo, hi)
        assert used == want, (loc, scale, used, want)
        assert mean == pytest.approx(t[:want].mean(), rel=1e-14)
    # constant values stop at the first batch end past min_k
    assert orl.adaptive_trace(np.full(100, 3.0), 30, 100) == (3.0, 32)


def test_beta_cov_is_the_gls_covariance():
    """sigma_e^2 C^-1 restricted to the fixed effects equals the GLS
    covariance (X'V^-1 X)^-1 with V = vu Z A Z' + ve I, A = Q^-1 (Henderson),
    with a record lacking a level and two records sharing one."""
    X, aid, y, Q, _ = small_problem(n=80, seed=21)
    aid = aid.copy()
    aid[3], aid[9] = -1, aid[8]
    vu, ve = 0.6, 1.4
    Ci, se = orl.beta_cov(X, aid, y, Q, vu, ve)
    Z = orl.design_w(np.zeros((X.shape[0], 0)), aid, Q.shape[0])
    V = vu * Z @ np.linalg.inv(Q) @ Z.T + ve * np.eye(X.shape[0])
    gls = np.linalg.inv(X.T @ np.linalg.solve(V, X))
    np.testing.assert_allclose(ve * Ci, gls, rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(se, np.sqrt(np.diag(gls)), rtol=1e-8)
