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


def test_mv_solve_matches_oracle(cb, ctx):
    Y, X, G = mv_problem(n=300)
    Gi = np.linalg.inv(G)
    Vg = np.array([[0.9, 0.3, 0.0], [0.3, 0.7, 0.1], [0.0, 0.1, 0.5]])
    Ve = np.array([[1.1, 0.1, 0.2], [0.1, 0.9, 0.0], [0.2, 0.0, 1.0]])
    beta, u = multitrait.solveMultiTrait(Y, X, Gi, Vg, Ve)
    br, ur = orl.mv_solve(Y, X, Gi, Vg, Ve)
    np.testing.assert_allclose(beta, br, atol=1e-9 * np.abs(br).max())
    np.testing.assert_allclose(u, ur, atol=1e-9 * np.abs(ur).max())


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
