"""GPU parity of the MME operator, block PCG (FP64 plain and mixed FP32 +
FP64 refinement) and AI/EM-REML against the dense FP64 oracle on the same
inputs and the same probe vectors.

Tolerances (the accuracy bar of the mixed-precision path): EBV error
max|du| / max|u| <= 1e-6; variance components relative error <= 1e-4; the
operator itself to 1e-12 relative.
"""
import numpy as np
import pytest

from cosmicblup_amd import _lib, design, mme, synth
from oracle import grm as og
from oracle import reml as orl

pytestmark = pytest.mark.gpu

EBV_TOL = 1e-6
VC_TOL = 1e-4


def gblup_problem(n=400, m=2000, ngroups=8, seed=11, h2=0.3, missing_records=0):
    x, _ = synth.simulate_dosages(n, m, maf=(0.05, 0.5), seed=seed)
    bed = synth.encode_bed(x)
    G, _, _ = og.grm_vanraden(bed, n, m)
    Q = og.invert_psd_matrix(G)
    y, cg = synth.simulate_phenotype(x, n_qtl=50, h2=h2, n_groups=ngroups, seed=seed + 1)
    aid = np.arange(n, dtype=np.int32)
    if missing_records:
        keep = np.ones(n, bool)
        keep[np.random.default_rng(seed).choice(n, missing_records, replace=False)] = False
        y, aid = y[keep], aid[keep]
        cg = [c for c, k in zip(cg, keep) if k]
    fd = design.build_fixed_design([[c] for c in cg])
    return fd, aid, y, Q


def ebv_err(u, u_ref):
    return np.abs(u - u_ref).max() / np.abs(u_ref).max()


def test_operator_matches_dense(cb, ctx):
    fd, aid, y, Q = gblup_problem()
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    C, rhs, _ = orl.mme_dense(design.dense_x(fd), aid, y, Q, 2.3)
    assert np.allclose(M.rhs(), rhs, rtol=1e-14, atol=1e-12)
    Xb = np.random.default_rng(0).standard_normal((M.dim, 70))
    for r in (1, 2, 3, 70):
        Y = M.apply(2.3, Xb[:, :r])
        ref = C @ Xb[:, :r]
        assert np.abs(Y - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("r", [20, 65])
def test_f16_operator_solves_match_fp64_at_several_splits(cb, ctx, r):
    """The FP16-halves product at q = 2999: 24 row tiles (the last partial),
    a last 16-row strip half padding, several k splits with a short last
    one, and a padded 16-column tile (r = 20) or five tiles (r = 65). Its
    solves must match FP64 ones; a wrong fragment, split or scale would
    leave errors far above the FP64 tolerance."""
    n, rng = 2999, np.random.default_rng(29)
    Z = rng.standard_normal((n, 400))
    G = Z @ Z.T / 400 + 0.1 * np.eye(n)
    Q = np.linalg.inv(G)
    Q = (Q + Q.T) / 2
    fd = design.build_fixed_design([[f"g{i % 5}"] for i in range(n)])
    y = rng.standard_normal(n) + np.arange(n) % 5
    M = mme.MixedModel(fd, np.arange(n, dtype=np.int32), y, n)
    M.set_qinv_dense(Q)
    B = np.zeros((M.dim, r))
    B[:, 0] = M.rhs()
    B[M.p:, 1:] = orl.probes(M.q, r - 1, seed=5)
    X64, _, _ = M.solve(0.8, B, mme.solve_options(precision=_lib.CB_PREC_FP64, tol=1e-12))
    X16, _, rel = M.solve(0.8, B, mme.solve_options(precision=_lib.CB_PREC_MIXED_F16, tol=1e-11))
    assert rel < 1e-11
    for c in range(r):
        assert ebv_err(X16[M.p:, c], X64[M.p:, c]) < 1e-8, c


@pytest.mark.parametrize("precision", [_lib.CB_PREC_FP64, _lib.CB_PREC_MIXED, _lib.CB_PREC_MIXED_F16])
@pytest.mark.parametrize("precond", [_lib.CB_PRECOND_JACOBI, _lib.CB_PRECOND_BLOCK_JACOBI])
def test_pcg_solves_match_dense(cb, ctx, precision, precond):
    fd, aid, y, Q = gblup_problem(missing_records=17)
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    lam = 1.7
    C, rhs, _ = orl.mme_dense(design.dense_x(fd), aid, y, Q, lam)
    B = np.zeros((M.dim, 5))
    B[:, 0] = rhs
    B[M.p:, 1:] = orl.probes(M.q, 4, seed=3)
    # the mixed path tests the true FP64 residual, whose rounding floor sits
    # near 1e-12 here; the FP64 path tests the recursive residual
    tol = 1e-12 if precision == _lib.CB_PREC_FP64 else 1e-10
    X, it, rel = M.solve(lam, B, mme.solve_options(precision=precision, precond=precond, tol=tol))
    ref = np.linalg.solve(C, B)
    assert rel < tol
    for c in range(5):
        assert ebv_err(X[M.p:, c], ref[M.p:, c]) < EBV_TOL
    # fixed effects are not identifiable beyond X b: compare fitted values
    Xd = design.dense_x(fd)
    np.testing.assert_allclose(Xd @ X[:M.p, 0], Xd @ ref[:M.p, 0], atol=1e-8)


def test_plain_unbatched_equals_batched(cb, ctx):
    fd, aid, y, Q = gblup_problem(n=300, m=1500)
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    B = np.zeros((M.dim, 3))
    B[:, 0] = M.rhs()
    B[M.p:, 1:] = orl.probes(M.q, 2, seed=8)
    X1, _, _ = M.solve(1.1, B, mme.solve_options(precision=_lib.CB_PREC_FP64, batch_rhs=True, tol=1e-12))
    X2, _, _ = M.solve(1.1, B, mme.solve_options(precision=_lib.CB_PREC_FP64, batch_rhs=False, tol=1e-12))
    assert np.abs(X1 - X2).max() < 1e-9 * np.abs(X1).max()


def test_absolute_stop_rule(cb, ctx):
    fd, aid, y, Q = gblup_problem(n=200, m=800)
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    b = M.rhs()
    X, it, res = M.solve(2.0, b, mme.solve_options(precision=_lib.CB_PREC_FP64,
                                                  stop_rule=_lib.CB_STOP_ABSOLUTE, tol=1e-6))
    C, _, _ = orl.mme_dense(design.dense_x(fd), aid, y, Q, 2.0)
    assert np.linalg.norm(C @ X - b) < 1e-6


def test_iteration_cap_reports_not_converged(cb, ctx):
    fd, aid, y, Q = gblup_problem(n=200, m=800)
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    with pytest.raises(cb.NotConverged):
        M.solve(2.0, M.rhs(), mme.solve_options(precision=_lib.CB_PREC_FP64, tol=1e-14, max_iter=2))


def test_probes_match_oracle(cb, ctx):
    fd, aid, y, Q = gblup_problem(n=150, m=600)
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    R = mme.AIREML(M, mme.reml_options(n_probes=9, seed=1234))
    assert np.array_equal(R.probes(), orl.probes(M.q, 9, 1234))


@pytest.mark.parametrize("precision", [_lib.CB_PREC_FP64, _lib.CB_PREC_MIXED, _lib.CB_PREC_MIXED_F16])
def test_score_and_ai_match_oracle(cb, ctx, precision):
    fd, aid, y, Q = gblup_problem()
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    opts = mme.reml_options(n_probes=16, seed=77,
                            solve=mme.solve_options(precision=precision,
                                                    tol=1e-12 if precision == _lib.CB_PREC_FP64 else 1e-10))
    R = mme.AIREML(M, opts)
    g, ai, tr, sol = R.score(0.5, 0.8)
    ref = orl.reml_eval(design.dense_x(fd), aid, y, Q, 0.5, 0.8, orl.probes(M.q, 16, 77))
    np.testing.assert_allclose(tr, ref["tr"], rtol=1e-8)
    np.testing.assert_allclose(g, ref["grad"], rtol=1e-7, atol=1e-9 * np.abs(ref["grad"]).max())
    np.testing.assert_allclose(ai, ref["ai"], rtol=1e-7)
    assert ebv_err(sol[M.p:], ref["sol"][M.p:]) < EBV_TOL


@pytest.mark.parametrize("qkind", ["dense", "csr"])
def test_probe_chunks_match_one_block(cb, ctx, qkind):
    """A block too narrow for y and every probe takes the probes in chunks,
    regenerating them and their Q z at each evaluation: the same score, AI
    and trace as one block, to the solver's tolerance, against the oracle
    on the same probes; probes copied out chunk by chunk are the same."""
    if qkind == "dense":
        fd, aid, y, Q = gblup_problem()
        M = mme.MixedModel(fd, aid, y, Q.shape[0])
        M.set_qinv_dense(Q)
    else:
        from cosmicblup_amd.pedigree import invertA_henderson

        dam, sire = synth.simulate_pedigree(4, 100, seed=21)
        A = invertA_henderson(dam, sire)
        rng = np.random.default_rng(8)
        aid = np.sort(rng.choice(dam.size, 300, replace=False)).astype(np.int32)
        y = rng.standard_normal(aid.size) + 3.0
        fd = design.build_fixed_design([[f"g{int(v)}"] for v in rng.integers(0, 5, aid.size)])
        M = mme.MixedModel(fd, aid, y, dam.size)
        M.set_qinv_csr(A.ptr, A.col, A.val)
    solve = mme.solve_options(precision=_lib.CB_PREC_FP64, tol=1e-12)
    one = mme.AIREML(M, mme.reml_options(n_probes=21, seed=5, solve=solve))
    g1, ai1, tr1, sol1 = one.score(0.6, 0.9)
    for bc in (2, 5, 8):
        ch = mme.AIREML(M, mme.reml_options(n_probes=21, seed=5, solve=solve, block_cols=bc))
        assert np.array_equal(ch.probes(), one.probes())
        g, ai, tr, sol = ch.score(0.6, 0.9)
        np.testing.assert_allclose(tr, tr1, rtol=1e-9)
        np.testing.assert_allclose(g, g1, rtol=1e-8, atol=1e-10 * np.abs(g1).max())
        np.testing.assert_allclose(ai, ai1, rtol=1e-9)
        assert ebv_err(sol[M.p:], sol1[M.p:]) < 1e-9
    if qkind == "dense":
        ref = orl.reml_eval(design.dense_x(fd), aid, y, Q, 0.6, 0.9, orl.probes(M.q, 21, 5))
        np.testing.assert_allclose(tr1, ref["tr"], rtol=1e-8)
        np.testing.assert_allclose(g1, ref["grad"], rtol=1e-7, atol=1e-9 * np.abs(ref["grad"]).max())


def test_exact_trace_in_chunks_and_run(cb, ctx):
    """Exact trace (q unit probes) through 7-column chunks equals the
    one-block exact trace; a chunked AI-REML run reaches the same estimates."""
    fd, aid, y, Q = gblup_problem(n=120, m=600, ngroups=3)
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    solve = mme.solve_options(precision=_lib.CB_PREC_FP64, tol=1e-12)
    g1, _, tr1, _ = mme.AIREML(M, mme.reml_options(trace_exact=True, solve=solve)).score(0.45, 0.7, want_ai=False)
    g2, _, tr2, _ = mme.AIREML(M, mme.reml_options(trace_exact=True, solve=solve, block_cols=7)).score(
        0.45, 0.7, want_ai=False)
    np.testing.assert_allclose(tr2, tr1, rtol=1e-9)
    np.testing.assert_allclose(g2, g1, rtol=1e-8, atol=1e-10 * np.abs(g1).max())
    runs = []
    for bc in (0, 6):
        R = mme.AIREML(M, mme.reml_options(n_probes=16, seed=3, solve=solve, block_cols=bc))
        R.setInitialVars(0.5, 0.5)
        R.solve()
        runs.append(R)
    # the one-block run starts each evaluation's solves from the previous
    # solutions, the chunked one from zero: equal to the solves' tolerance
    assert runs[0].iterations == runs[1].iterations
    np.testing.assert_allclose([runs[1].var_u, runs[1].var_e], [runs[0].var_u, runs[0].var_e], rtol=1e-6,
                               atol=1e-9)
    assert ebv_err(runs[1].solution[M.p:], runs[0].solution[M.p:]) < EBV_TOL


def test_exact_trace_score_matches_fd_of_logL(cb, ctx):
    fd, aid, y, Q = gblup_problem(n=120, m=600, ngroups=3)
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    R = mme.AIREML(M, mme.reml_options(trace_exact=True, solve=mme.solve_options(tol=1e-10)))
    vu, ve = 0.45, 0.7
    g, _, _, _ = R.score(vu, ve, want_ai=False)
    X = design.dense_x(fd)
    h = 1e-4
    fd_u = (orl.logL(X, aid, y, Q, vu * (1 + h), ve) - orl.logL(X, aid, y, Q, vu * (1 - h), ve)) / (2 * h * vu)
    fd_e = (orl.logL(X, aid, y, Q, vu, ve * (1 + h)) - orl.logL(X, aid, y, Q, vu, ve * (1 - h))) / (2 * h * ve)
    np.testing.assert_allclose(g, [fd_u, fd_e], rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("algorithm", [_lib.CB_VCE_AI, _lib.CB_VCE_EM])
def test_reml_run_matches_oracle(cb, ctx, algorithm):
    fd, aid, y, Q = gblup_problem()
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    k, seed = 32, 5
    opts = mme.reml_options(algorithm=algorithm, n_probes=k, seed=seed, max_iter=20,
                            solve=mme.solve_options(tol=1e-10))
    R = mme.AIREML(M, opts)
    R.solve()
    ref = orl.reml_run(design.dense_x(fd), aid, y, Q, algorithm="ai" if algorithm == _lib.CB_VCE_AI else "em",
                       max_iter=20, Zp=orl.probes(M.q, k, seed))
    assert R.iterations == len(ref["hist"])
    np.testing.assert_allclose([R.var_u, R.var_e], [ref["vu"], ref["ve"]], rtol=VC_TOL)
    assert ebv_err(R.solution[M.p:], ref["sol"][M.p:]) < EBV_TOL
    # calculate_SE from the last AI matrix
    np.testing.assert_allclose(R.ai_last, ref["ai_last"], rtol=1e-4)
    np.testing.assert_allclose([R.getSEVarsU()[0], R.getSEVarE()], ref["se"], rtol=1e-4)


@pytest.mark.parametrize("which", ["dense", "csr"])
def test_logl_matches_oracle(cb, ctx, which):
    """calc_logL_internal (vce.cpp:3533-3588) through a device Cholesky."""
    if which == "dense":
        fd, aid, y, Q = gblup_problem(n=300, m=1500, missing_records=11)
        M = mme.MixedModel(fd, aid, y, Q.shape[0])
        M.set_qinv_dense(Q)
    else:
        from cosmicblup_amd.pedigree import invertA_henderson

        dam, sire = synth.simulate_pedigree(5, 80, seed=9)
        A = invertA_henderson(dam, sire)
        rng = np.random.default_rng(1)
        aid = np.sort(rng.choice(dam.size, 300, replace=False)).astype(np.int32)
        y = rng.standard_normal(aid.size) + 2.0
        fd = design.build_fixed_design([[f"g{int(v)}"] for v in rng.integers(0, 6, aid.size)])
        M = mme.MixedModel(fd, aid, y, dam.size)
        M.set_qinv_csr(A.ptr, A.col, A.val)
        Q = A.toDense()
    R = mme.AIREML(M)
    X = design.dense_x(fd)
    for vu, ve in ((0.5, 0.8), (1.3, 0.4)):
        np.testing.assert_allclose(R.logL(vu, ve), orl.logL(X, aid, y, Q, vu, ve), rtol=1e-10)


def test_pblup_native_ainv_with_inbreeding_mixed_vs_dense(cb, ctx):
    """PBLUP on the library's own A^-1 (inbred pedigree): mixed and FP64 PCG
    against the dense solve, and REML against the oracle on the same A^-1."""
    from cosmicblup_amd.pedigree import invertA_henderson

    dam, sire = synth.simulate_pedigree(6, 150, seed=12)
    A = invertA_henderson(dam, sire)
    assert A.F.max() > 0
    rng = np.random.default_rng(5)
    rec = np.sort(rng.choice(dam.size, 600, replace=False)).astype(np.int32)
    y = rng.standard_normal(rec.size) + 5.0
    fd = design.build_fixed_design([[f"g{int(v):02d}"] for v in rng.integers(0, 12, rec.size)])
    M = mme.MixedModel(fd, rec, y, dam.size)
    M.set_qinv_csr(A.ptr, A.col, A.val)
    Ad = A.toDense()
    C, rhs, _ = orl.mme_dense(design.dense_x(fd), rec, y, Ad, 2.0)
    ref = np.linalg.solve(C, rhs)
    for prec in (_lib.CB_PREC_FP64, _lib.CB_PREC_MIXED, _lib.CB_PREC_MIXED_F16):
        X, _, _ = M.solve(2.0, rhs, mme.solve_options(precision=prec, tol=1e-12 if prec == 0 else 1e-10))
        assert ebv_err(X[M.p:], ref[M.p:]) < EBV_TOL
    R = mme.AIREML(M, mme.reml_options(n_probes=16, seed=4, solve=mme.solve_options(tol=1e-10)))
    R.solve()
    o = orl.reml_run(design.dense_x(fd), rec, y, Ad, Zp=orl.probes(dam.size, 16, 4))
    np.testing.assert_allclose([R.var_u, R.var_e], [o["vu"], o["ve"]], rtol=VC_TOL)


def test_pblup_mean_row_of_several_chunks(cb, ctx):
    """PBLUP with 8000 records: the mean row of W'W has more than one chunk
    of 4096 nonzeros, so the CG iteration finishes it from its chunk partials
    inside the dot and update kernels, and the 40 group rows are one chunk
    each. FP64 and mixed solutions against a sparse direct solve of the same
    MME assembled on the host with scipy."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    from cosmicblup_amd.pedigree import invertA_henderson

    dam, sire = synth.simulate_pedigree(3, 4000, unknown=0.05, seed=31)
    q = dam.size
    A = invertA_henderson(dam, sire)
    rng = np.random.default_rng(32)
    rec = np.sort(rng.choice(q, 8000, replace=False)).astype(np.int32)
    y = rng.standard_normal(rec.size) + 4.0
    fd = design.build_fixed_design([[f"g{int(v):02d}"] for v in rng.integers(0, 40, rec.size)])
    lam = 2.5
    X = sp.csr_matrix((fd.val, fd.col, fd.ptr), shape=(rec.size, fd.p))
    Zi = sp.csr_matrix((np.ones(rec.size), rec.astype(np.int64), np.arange(rec.size + 1)), shape=(rec.size, q))
    W = sp.hstack([X, Zi]).tocsr()
    Ai = sp.csr_matrix((A.val, A.col, A.ptr), shape=(q, q))
    Cm = (W.T @ W + sp.block_diag([sp.csr_matrix((fd.p, fd.p)), lam * Ai])).tocsc()
    b = W.T @ y
    ref = spl.spsolve(Cm, b)
    M = mme.MixedModel(fd, rec, y, q, ctx=ctx)
    M.set_qinv_csr(A.ptr, A.col, A.val)
    for prec in (_lib.CB_PREC_FP64, _lib.CB_PREC_MIXED):
        x, it, rel = M.solve(lam, b, mme.solve_options(precision=prec, tol=1e-12, max_iter=5000))
        assert rel < 1e-12, (prec, it, rel)
        assert np.linalg.norm(Cm @ x - b) / np.linalg.norm(b) < 1e-11
        assert ebv_err(x[M.p:], ref[M.p:]) < EBV_TOL


def test_dense_q_mean_row_of_several_chunks(cb, ctx):
    """Dense Q with 4600 records: the mean row of W'W (4611 nonzeros) has two
    chunks of 4096, formed in the same pass as the short rows and finished
    inside the p.q dot; FP64 and mixed solutions against a dense solve."""
    rng = np.random.default_rng(44)
    n = 4600
    B = rng.standard_normal((n, 40))
    Q = B @ B.T / 40.0 + np.eye(n)
    y = rng.standard_normal(n) + 2.0
    fd = design.build_fixed_design([[f"g{int(v)}"] for v in rng.integers(0, 10, n)])
    aid = np.arange(n, dtype=np.int32)
    M = mme.MixedModel(fd, aid, y, n)
    M.set_qinv_dense(Q)
    C, rhs, _ = orl.mme_dense(design.dense_x(fd), aid, y, Q, 1.7)
    ref = np.linalg.solve(C, rhs)
    for prec in (_lib.CB_PREC_FP64, _lib.CB_PREC_MIXED):
        X, _, _ = M.solve(1.7, rhs, mme.solve_options(precision=prec, tol=1e-11))
        assert ebv_err(X[M.p:], ref[M.p:]) < EBV_TOL
        assert np.abs(X[:M.p] - ref[:M.p]).max() / np.abs(ref[:M.p]).max() < 1e-6


def test_pblup_csr_ainv(cb, ctx):
    """PBLUP with Henderson A^-1 as CSR (SparseMatrixAdapter path)."""
    rng = np.random.default_rng(2)
    nan = 600
    sire = np.full(nan, -1)
    dam = np.full(nan, -1)
    for i in range(100, nan):
        sire[i] = rng.integers(0, 50)
        dam[i] = rng.integers(50, 100)
    Ai = orl.henderson_ainv(sire, dam)
    nz = np.nonzero(Ai)
    ptr = np.zeros(nan + 1, dtype=np.int64)
    np.add.at(ptr, nz[0] + 1, 1)
    ptr = np.cumsum(ptr)
    col = nz[1].astype(np.int32)
    val = Ai[nz]
    rec = rng.choice(nan, 450, replace=False).astype(np.int32)
    y = rng.standard_normal(rec.size) + 3.0
    cats = [[f"g{int(v)}"] for v in rng.integers(0, 5, rec.size)]
    fd = design.build_fixed_design(cats)
    M = mme.MixedModel(fd, rec, y, nan)
    M.set_qinv_csr(ptr, col, val)
    C, rhs, _ = orl.mme_dense(design.dense_x(fd), rec, y, Ai, 1.5)
    X, _, _ = M.solve(1.5, rhs, mme.solve_options(tol=1e-10))
    ref = np.linalg.solve(C, rhs)
    assert ebv_err(X[M.p:], ref[M.p:]) < EBV_TOL
    opts = mme.reml_options(n_probes=16, seed=3, solve=mme.solve_options(tol=1e-10))
    R = mme.AIREML(M, opts)
    R.solve()
    o = orl.reml_run(design.dense_x(fd), rec, y, Ai, Zp=orl.probes(nan, 16, 3))
    np.testing.assert_allclose([R.var_u, R.var_e], [o["vu"], o["ve"]], rtol=VC_TOL)


def test_dense_exact_score_and_run_match_oracle(cb, ctx):
    """run_dense_exact_step (vce.cpp:2859-3195) on the device vs the oracle."""
    fd, aid, y, Q = gblup_problem(n=300, m=1500, missing_records=9)
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    X = design.dense_x(fd)
    opts = mme.reml_options(algorithm=_lib.CB_VCE_AI_EXACT, max_iter=30, tol=1e-6)
    R = mme.AIREML(M, opts)
    g, ai, _, sol = R.score(0.6, 0.9)
    t = orl.dense_exact_eval(X, aid, y, Q, 0.6, 0.9)
    np.testing.assert_allclose(g, t["grad"], rtol=1e-8, atol=1e-10 * np.abs(t["grad"]).max())
    np.testing.assert_allclose(ai, t["ai"], rtol=1e-9)
    assert ebv_err(sol[M.p:], t["sol"][M.p:]) < 1e-10
    R.solve()
    o = orl.dense_exact_run(X, aid, y, Q, max_iter=30, tol=1e-6)
    assert R.iterations == len(o["hist"]) and R.converged == o["converged"]
    # the same Newton path; its last step (< tol = 1e-6) carries the rounding
    # differences of the two factorisations
    np.testing.assert_allclose([R.var_u, R.var_e], [o["vu"], o["ve"]], rtol=1e-6)
    np.testing.assert_allclose([R.getSEVarsU()[0], R.getSEVarE()], o["se"], rtol=1e-5)
    assert ebv_err(R.solution[M.p:], o["sol"][M.p:]) < 1e-6


@pytest.mark.parametrize("precision", [_lib.CB_PREC_FP64, _lib.CB_PREC_MIXED, _lib.CB_PREC_MIXED_F16])
@pytest.mark.parametrize("qkind", ["dense", "csr"])
def test_graph_replay_equals_plain_launches(cb, ctx, precision, qkind):
    """PCG iteration blocks replayed as HIP graphs give the same iterates,
    bit for bit, and the same iteration counts as one launch per kernel.
    Each solve replays its graph many times, with host checks (and, mixed,
    FP64 residual replacements) between replays; solves alternate between
    the two modes so that every graph is built afresh."""
    if qkind == "dense":
        fd, aid, y, Q = gblup_problem(missing_records=17)
        M = mme.MixedModel(fd, aid, y, Q.shape[0])
        M.set_qinv_dense(Q)
        B = np.zeros((M.dim, 5))
        B[:, 0] = M.rhs()
        B[M.p:, 1:] = orl.probes(M.q, 4, seed=3)
    else:
        from cosmicblup_amd.pedigree import invertA_henderson

        dam, sire = synth.simulate_pedigree(6, 150, seed=12)
        A = invertA_henderson(dam, sire)
        rng = np.random.default_rng(5)
        rec = np.sort(rng.choice(dam.size, 600, replace=False)).astype(np.int32)
        y = rng.standard_normal(rec.size) + 5.0
        fd = design.build_fixed_design([[f"g{int(v):02d}"] for v in rng.integers(0, 12, rec.size)])
        M = mme.MixedModel(fd, rec, y, dam.size)
        M.set_qinv_csr(A.ptr, A.col, A.val)
        B = M.rhs()
    opts = mme.solve_options(precision=precision, tol=1e-12 if precision == _lib.CB_PREC_FP64 else 1e-10)
    runs = []
    try:
        for graphs in (False, True, False, True):
            ctx.set_graphs(graphs)
            runs.append(M.solve(1.7, B, opts))
    finally:
        ctx.set_graphs(False)
    X0, it0, rel0 = runs[0]
    assert it0 >= 24  # several graph replays
    for X, it, rel in runs[1:]:
        assert it == it0
        assert rel == rel0
        assert np.array_equal(X, X0)


def test_contexts_on_threads_match_serial(cb, ctx):
    """One context per host thread, each with its own models: the same
    estimates, bit for bit, as the same calls made one after another (the
    library's host copies share one staging buffer; its device work shares
    nothing between contexts)."""
    from concurrent.futures import ThreadPoolExecutor

    ctxs = [_lib.Context(0) for _ in range(3)]
    models = []
    for w, c in enumerate(ctxs):
        for j in range(2):
            fd, aid, y, Q = gblup_problem(n=250 + 90 * (2 * w + j), m=1200, seed=40 + 2 * w + j)
            M = mme.MixedModel(fd, aid, y, Q.shape[0], ctx=c)
            M.set_qinv_dense(Q)
            models.append((w, M))
    opts = mme.reml_options(n_probes=12, seed=9, solve=mme.solve_options(tol=1e-10))

    def run(M):
        R = mme.AIREML(M, opts)
        R.setInitialVars(0.5, 0.5)
        R.solve()
        X, it, _ = M.solve(1.3, M.rhs(), mme.solve_options(precision=_lib.CB_PREC_FP64, tol=1e-12))
        return R.iterations, R.var_u, R.var_e, R.solution.tobytes(), X.tobytes(), it

    serial = [run(M) for _, M in models]

    def worker(w):
        return [(i, run(M)) for i, (wk, M) in enumerate(models) if wk == w]

    with ThreadPoolExecutor(len(ctxs)) as ex:
        par = dict(sum(ex.map(worker, range(len(ctxs))), []))
    for i, ref in enumerate(serial):
        assert par[i] == ref


@pytest.mark.parametrize("precision", [_lib.CB_PREC_FP64, _lib.CB_PREC_MIXED, _lib.CB_PREC_MIXED_F16])
def test_device_resident_solve_equals_host_solve(cb, ctx, precision):
    """cb_mme_solve_device on torch-owned blocks with padded leading
    dimensions gives the host-array solve's iterates bit for bit, and leaves
    the padding rows of X alone."""
    import torch

    fd, aid, y, Q = gblup_problem(n=300, m=1500, missing_records=9)
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    B = np.zeros((M.dim, 3))
    B[:, 0] = M.rhs()
    B[M.p:, 1:] = orl.probes(M.q, 2, seed=4)
    opts = mme.solve_options(precision=precision, tol=1e-12 if precision == _lib.CB_PREC_FP64 else 1e-10)
    X, it, rel = M.solve(1.4, B, opts)
    ldb, ldx = M.dim + 5, M.dim + 11
    dB = torch.zeros((3, ldb), dtype=torch.float64, device="cuda")
    dB[:, :M.dim] = torch.from_numpy(B.T.copy()).cuda()
    dX = torch.full((3, ldx), -7.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    it_d, rel_d = M.solve_device(1.4, dB.data_ptr(), ldb, dX.data_ptr(), ldx, 3, opts)
    Xd = dX.cpu().numpy()
    assert it_d == it and rel_d == rel
    assert np.array_equal(Xd[:, :M.dim].T, X)
    assert np.all(Xd[:, M.dim:] == -7.0)


@pytest.mark.parametrize("precision", [_lib.CB_PREC_FP64, _lib.CB_PREC_MIXED])
@pytest.mark.parametrize("qkind", ["dense", "csr"])
def test_beta_cov_matches_oracle(cb, ctx, precision, qkind):
    """GSSolve's fixed-effect SEs (mixed_model_solver.cpp:346-417) from the
    device's unit-column block solves, 300 fixed effects (two blocks of
    right-hand sides), against the dense oracle."""
    rng = np.random.default_rng(41)
    n = 1500
    if qkind == "dense":
        x, _ = synth.simulate_dosages(n, 4000, maf=(0.05, 0.5), seed=41)
        G, _, _ = og.grm_vanraden(synth.encode_bed(x), n, 4000)
        Q = og.invert_psd_matrix(G)
    else:
        from cosmicblup_amd.pedigree import invertA_henderson

        dam, sire = synth.simulate_pedigree(3, n // 3, unknown=0.05, seed=41)
        A = invertA_henderson(dam, sire)
        Q = A.toDense()
    grp = rng.integers(0, 300, n)
    fd = design.build_fixed_design([[f"g{int(c):03d}"] for c in grp])
    aid = np.arange(n, dtype=np.int32)
    y = rng.standard_normal(n) + grp * 0.01
    M = mme.MixedModel(fd, aid, y, n)
    if qkind == "dense":
        M.set_qinv_dense(Q)
    else:
        M.set_qinv_csr(A.ptr, A.col, A.val)
    vu, ve = 0.4, 0.9
    tol = 1e-12 if precision == _lib.CB_PREC_FP64 else 1e-10
    cinv, se = M.beta_cov(ve / vu, ve, mme.solve_options(precision=precision, tol=tol))
    cref, sref = orl.beta_cov(design.dense_x(fd), aid, y, Q, vu, ve)
    assert M.p > 256
    np.testing.assert_allclose(se, sref, rtol=1e-7)
    np.testing.assert_allclose(cinv, cref, atol=1e-8 * np.abs(cref).max())
