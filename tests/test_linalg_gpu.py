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
    lo, hi = linalg.invert_psd_device(d.data_ptr(), 128, 128, d.data_ptr(), 128)
    ev = np.linalg.eigvalsh(M)
    assert abs(lo - ev[0]) < 1e-12 * ev[-1] and abs(hi - ev[-1]) < 1e-12 * ev[-1]
    np.testing.assert_allclose(d.cpu().numpy(), og.invert_psd_matrix(M), rtol=0, atol=1e-10 * np.abs(np.linalg.inv(M)).max())


@pytest.mark.parametrize("lam_min", [2e-6, 5e-7])
def test_floor_inactive_cholesky_path_and_active_floor(cb, ctx, lam_min):
    """A symmetric matrix with prescribed eigenvalues: its smallest just above
    the 1e-6 floor (the library inverts by Cholesky, the floor being
    inactive) or just below it (the eigensolve path floors it). Both match
    the oracle's eigen-bent inverse."""
    n = 400
    rng = np.random.default_rng(5)
    Qm, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.linspace(0.05, 3.0, n)
    lam[0] = lam_min
    M = (Qm * lam) @ Qm.T
    M = 0.5 * (M + M.T)
    got = linalg.invert_psd_matrix(M)
    ref = og.invert_psd_matrix(M)
    assert np.abs(got - ref).max() <= 1e-8 * np.abs(ref).max()
    if lam_min > 1e-6:  # the Cholesky inverse is mirrored, so exactly symmetric
        assert np.array_equal(got, got.T)
    top = np.linalg.eigvalsh(got).max()
    assert abs(top - 1.0 / max(lam_min, 1e-6)) < 1e-6 * top
