"""GPU parity of the eigen-bent inverse (matrix_io.cpp:488-501) against the
oracle's numpy/LAPACK restatement."""
import numpy as np
import pytest

from cosmicblup_amd import linalg, synth
from oracle import grm as og

pytestmark = pytest.mark.gpu


def test_bent_inverse_of_singular_grm_matches_oracle(cb, ctx):
    n, m = 300, 1500
    x, _ = synth.simulate_dosages(n, m, maf=(0.05, 0.5), seed=21)
    G, _, _ = og.grm_vanraden(synth.encode_bed(x), n, m)  # centred: one eigenvalue ~ 0
    got = linalg.invert_psd_matrix(G)
    ref = og.invert_psd_matrix(G)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()
    # the floored direction: largest eigenvalue of the inverse is 1 / 1e-6
    assert abs(np.linalg.eigvalsh(got).max() - 1e6) < 1e-3


def test_inverse_with_ridge_of_spd_matrix(cb, ctx):
    rng = np.random.default_rng(3)
    B = rng.standard_normal((200, 260))
    M = B @ B.T / 260
    got = linalg.invert_psd_matrix(M, ridge=0.01)
    np.testing.assert_allclose(got @ (M + 0.01 * np.eye(200)), np.eye(200), atol=1e-9)
    np.testing.assert_allclose(got, og.invert_psd_matrix(M, ridge=0.01), rtol=0, atol=1e-9 * np.abs(got).max())


def test_device_variant_in_place(cb, ctx):
    import torch

    rng = np.random.default_rng(4)
    B = rng.standard_normal((128, 150))
    M = B @ B.T / 150
    d = torch.from_numpy(M.copy()).cuda()
    torch.cuda.synchronize()
    lo, hi = linalg.invert_psd_device(d.data_ptr(), 128, 128, d.data_ptr(), 128, eigenvalues=True)
    ev = np.linalg.eigvalsh(M)
    assert abs(lo - ev[0]) < 1e-12 * ev[-1] and abs(hi - ev[-1]) < 1e-12 * ev[-1]
    np.testing.assert_allclose(d.cpu().numpy(), og.invert_psd_matrix(M), rtol=0, atol=1e-10 * np.abs(np.linalg.inv(M)).max())


@pytest.mark.parametrize("small", [(2e-6,), (5e-7,), (-3e-9,), (5e-7, 3e-7)])
def test_bent_inverse_paths_match_oracle(cb, ctx, small):
    """A symmetric matrix with prescribed eigenvalues around the 1e-6 floor:
    all above it (the floor is inactive: Cholesky inverse), one below it (one
    eigenpair by inverse iteration, floored by a rank-1 update, Cholesky
    inverse), two below it (the eigensolve). Each matches the oracle's
    eigen-bent inverse, and the floored directions sit at 1 / 1e-6."""
    n = 400
    rng = np.random.default_rng(5)
    Qm, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.linspace(0.05, 3.0, n)
    lam[: len(small)] = small
    M = (Qm * lam) @ Qm.T
    M = 0.5 * (M + M.T)
    got = linalg.invert_psd_matrix(M)
    ref = og.invert_psd_matrix(M)
    assert np.abs(got - ref).max() <= 1e-8 * np.abs(ref).max()
    if len(small) == 1:  # the Cholesky inverses are mirrored, so exactly symmetric
        assert np.array_equal(got, got.T)
    ev = np.sort(np.linalg.eigvalsh(got))[::-1]
    want = np.sort(1.0 / np.maximum(small, 1e-6))[::-1]
    np.testing.assert_allclose(ev[: len(small)], want, rtol=1e-6)
