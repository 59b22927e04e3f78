"""Multi-trait GBLUP on the device (MvEigenVce, eigen_vce.cpp:106-262;
solveMultiTrait, mixed_model_solver.cpp:425-524) against the oracle."""
import numpy as np
import pytest

from cosmicblup_amd import multitrait, synth
from oracle import grm as og
from oracle import reml as orl

pytestmark = pytest.mark.gpu


def mv_problem(n=400, t=3, seed=6):
    rng = np.random.default_rng(seed)
    x, _ = synth.simulate_dosages(n, 2000, maf=(0.05, 0.5), seed=seed)
    G, _, _ = og.grm_vanraden(synth.encode_bed(x), n, 2000)
    G = 0.95 * G + 0.05 * np.eye(n)
    Vg0 = np.array([[1.0, 0.5, -0.2], [0.5, 0.8, 0.1], [-0.2, 0.1, 0.6]])[:t, :t]
    Ve0 = np.array([[1.0, 0.2, 0.2], [0.2, 1.1, 0.2], [0.2, 0.2, 0.9]])[:t, :t]
    U = np.linalg.cholesky(G) @ rng.standard_normal((n, t)) @ np.linalg.cholesky(Vg0).T
    E = rng.standard_normal((n, t)) @ np.linalg.cholesky(Ve0).T
    X = np.column_stack([np.ones(n), (rng.integers(0, 3, n) == 1), (rng.integers(0, 3, n) == 2)]).astype(float)
    Y = X @ rng.standard_normal((3, t)) + U + E
    return Y, X, G


@pytest.mark.parametrize("t", [2, 3])
def test_mv_vce_matches_oracle(cb, ctx, t):
    Y, X, G = mv_problem(t=t)
    v = multitrait.MvEigenVce(multitrait.MvEigenVce.Options(max_iter=3000, tol=1e-7))
    v.prepare(Y, X, G)
    v.runNullModel()
    Vg, Ve, it, conv = orl.mv_eigen_vce(Y, X, G, max_iter=3000, tol=1e-7)
    assert v.converged == conv and abs(v.iterations - it) <= 1
    np.testing.assert_allclose(v.getVg(), Vg, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(v.getVe(), Ve, rtol=1e-5, atol=1e-7)


def test_mv_vce_iteration_cap_matches_oracle(cb, ctx):
    """The EM loop runs on the device: stopped by max_iter it reports the cap,
    not converged, with the estimates of that many EM steps."""
    Y, X, G = mv_problem(t=3)
    v = multitrait.MvEigenVce(multitrait.MvEigenVce.Options(max_iter=7, tol=1e-12))
    v.prepare(Y, X, G)
    v.runNullModel()
    Vg, Ve, it, conv = orl.mv_eigen_vce(Y, X, G, max_iter=7, tol=1e-12)
    assert (v.iterations, v.converged) == (it, conv) == (7, False)
    np.testing.assert_allclose(v.getVg(), Vg, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(v.getVe(), Ve, rtol=1e-9, atol=1e-12)


def test_mv_solve_matches_oracle(cb, ctx):
    Y, X, G = mv_problem(n=300)
    Gi = np.linalg.inv(G)
    Vg = np.array([[0.9, 0.3, 0.0], [0.3, 0.7, 0.1], [0.0, 0.1, 0.5]])
    Ve = np.array([[1.1, 0.1, 0.2], [0.1, 0.9, 0.0], [0.2, 0.0, 1.0]])
    beta, u = multitrait.solveMultiTrait(Y, X, Gi, Vg, Ve)
    br, ur = orl.mv_solve(Y, X, Gi, Vg, Ve)
    np.testing.assert_allclose(beta, br, atol=1e-9 * np.abs(br).max())
    np.testing.assert_allclose(u, ur, atol=1e-9 * np.abs(ur).max())


def bent_problem(n=1200, m=4000, seed=11):
    """A serving-workload-shaped 3-trait instance: centred VanRaden G (so its
    eigen-bent inverse has the 1e6 eigenvalue of the floored centring
    direction), 10 contemporary groups plus the mean, trait correlations
    as the batched workload draws them, and Vg / Ve from the oracle's EM
    stopped at its 500-iteration cap."""
    from cosmicblup_amd import design

    rng = np.random.default_rng(seed)
    x, _ = synth.simulate_dosages(n, m, maf=(0.05, 0.5), seed=seed)
    G, _, _ = og.grm_vanraden(synth.encode_bed(x), n, m)
    Gi = og.invert_psd_matrix(G)
    grp = rng.integers(0, 10, n)
    X = design.dense_x(design.build_fixed_design([[f"cg{int(c)}"] for c in grp]))
    Vg0 = np.array([[0.3, 0.1, -0.05], [0.1, 0.45, 0.2], [-0.05, 0.2, 0.2]])
    Ve0 = np.array([[0.7, 0.12, 0.1], [0.12, 0.55, 0.13], [0.1, 0.13, 0.8]])
    w, V = np.linalg.eigh(G)
    Gh = (V * np.sqrt(np.maximum(w, 0))) @ V.T
    U = Gh @ rng.standard_normal((n, 3)) @ np.linalg.cholesky(Vg0).T
    E = rng.standard_normal((n, 3)) @ np.linalg.cholesky(Ve0).T
    Y = X @ rng.standard_normal((X.shape[1], 3)) + U + E
    Vg, Ve, _, _ = orl.mv_eigen_vce(Y, X, G, max_iter=500)
    return Y, X, G, Gi, Vg, Ve


def test_mv_solve_bent_ginv_matches_oracle(cb, ctx):
    """solveMultiTrait on the eigen-bent G^-1 (largest eigenvalue 1e6) with
    EM-estimated covariances: beta and u within 1e-6 (relative to their
    largest entries) of the oracle's dense Kronecker solve, and the same bits
    on a second call."""
    Y, X, G, Gi, Vg, Ve = bent_problem()
    assert np.linalg.eigvalsh(Gi).max() > 1e5
    beta, u = multitrait.solveMultiTrait(Y, X, Gi, Vg, Ve)
    br, ur = orl.mv_solve(Y, X, Gi, Vg, Ve)
    assert np.abs(u - ur).max() <= 1e-6 * np.abs(ur).max()
    assert np.abs(beta - br).max() <= 1e-6 * np.abs(br).max()
    beta2, u2 = multitrait.solveMultiTrait(Y, X, Gi, Vg, Ve)
    assert np.array_equal(u, u2) and np.array_equal(beta, beta2)


def test_mv_solve_rejects_singular_vg(cb, ctx):
    Y, X, G = mv_problem(n=120)
    Vg = np.array([[1.0, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 0.5]])
    Ve = np.eye(3)
    from cosmicblup_amd import _lib

    with pytest.raises(_lib.CosmicError, match="Vg is not positive definite"):
        multitrait.solveMultiTrait(Y, X, np.linalg.inv(G), Vg, Ve)


def test_contexts_on_threads_match_serial(cb, ctx):
    """Multi-trait estimation and solve on three contexts driven by three
    host threads give the serial results bit for bit. These calls allocate
    and free device buffers while other contexts' work is queued: a buffer
    is released only behind its own context's work, so its memory is never
    handed to another stream early."""
    from concurrent.futures import ThreadPoolExecutor

    from cosmicblup_amd import _lib

    ctxs = [_lib.Context(0) for _ in range(3)]
    jobs = []
    for w, c in enumerate(ctxs):
        for j in range(2):
            Y, X, G = mv_problem(n=300 + 60 * (2 * w + j), seed=20 + 2 * w + j)
            jobs.append((w, c, Y, X, G, np.linalg.inv(G)))

    def run(c, Y, X, G, Gi):
        v = multitrait.MvEigenVce(multitrait.MvEigenVce.Options(max_iter=300, tol=1e-7), ctx=c)
        v.prepare(Y, X, G)
        v.runNullModel()
        beta, u = multitrait.solveMultiTrait(Y, X, Gi, v.getVg(), v.getVe(), ctx=c)
        return v.iterations, v.getVg().tobytes(), v.getVe().tobytes(), beta.tobytes(), u.tobytes()

    serial = [run(c, Y, X, G, Gi) for _, c, Y, X, G, Gi in jobs]

    def worker(w):
        return [(i, run(c, Y, X, G, Gi)) for i, (wk, c, Y, X, G, Gi) in enumerate(jobs) if wk == w]

    for _ in range(2):
        with ThreadPoolExecutor(len(ctxs)) as ex:
            par = dict(sum(ex.map(worker, range(len(ctxs))), []))
        for i, ref in enumerate(serial):
            assert par[i] == ref, i
