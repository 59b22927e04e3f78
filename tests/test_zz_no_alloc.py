"""The allocation-free device solve (cb_mme_solve_device_noalloc) against
cb_mme_solve_device on the same model and right-hand sides: two calls in a
row, the second with fewer columns, dense and CSR Qinv, FP64 and mixed
precision; the device's free memory is unchanged by the calls; sizes past
the scratch array are rejected."""
import numpy as np
import pytest

from cosmicblup_amd import _lib, design, mme, synth
from oracle import grm as og
from oracle import reml as orl

pytestmark = pytest.mark.gpu


def _model(qkind):
    n = 1800
    rng = np.random.default_rng(51)
    grp = rng.integers(0, 12, n)
    fd = design.build_fixed_design([[f"g{int(c):02d}"] for c in grp])
    y = rng.standard_normal(n) + grp * 0.1
    aid = np.arange(n, dtype=np.int32)
    M = mme.MixedModel(fd, aid, y, n)
    if qkind == "dense":
        x, _ = synth.simulate_dosages(n, 4000, maf=(0.05, 0.5), seed=51)
        G, _, _ = og.grm_vanraden(synth.encode_bed(x), n, 4000)
        M.set_qinv_dense(og.invert_psd_matrix(G))
    else:
        from cosmicblup_amd.pedigree import invertA_henderson

        dam, sire = synth.simulate_pedigree(3, n // 3, unknown=0.05, seed=51)
        A = invertA_henderson(dam, sire)
        M.set_qinv_csr(A.ptr, A.col, A.val)
    return M


@pytest.mark.parametrize("precision", [_lib.CB_PREC_FP64, _lib.CB_PREC_MIXED])
@pytest.mark.parametrize("qkind", ["dense", "csr"])
def test_noalloc_solve_matches_solve_device(cb, ctx, qkind, precision):
    import torch

    M = _model(qkind)
    dim = M.dim
    B = np.zeros((9, dim))
    B[0] = M.rhs()
    B[1:, M.p:] = orl.probes(M.q, 8, seed=5).T
    dB = torch.from_numpy(B).cuda()
    tol = 1e-12 if precision == _lib.CB_PREC_FP64 else 1e-10
    opts = mme.solve_options(precision=precision, tol=tol)
    lam = 1.3
    for r in (9, 4):  # two calls in a row, the second smaller
        dX_ref = torch.zeros((r, dim), dtype=torch.float64, device="cuda")
        dX = torch.full((r, dim), np.nan, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        it_ref, rel_ref = M.solve_device(lam, dB.data_ptr(), dim, dX_ref.data_ptr(), dim, r, opts)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        it, rel = M.solve_device_noalloc(lam, dB.data_ptr(), dim, dX.data_ptr(), dim, r, opts)
        free1 = torch.cuda.mem_get_info()[0]
        assert free1 == free0, "the allocation-free solve changed the device's free memory"
        x, xr = dX.cpu().numpy(), dX_ref.cpu().numpy()
        assert it == it_ref and rel == rel_ref
        assert np.array_equal(x, xr), np.abs(x - xr).max()
        assert rel < tol


def test_noalloc_rejects_sizes_past_its_scratch(cb, ctx):
    import torch

    M = _model("dense")
    r = 1000  # 1000 columns of 1813 unknowns: about 90 MB of scratch
    dB = torch.zeros((r, M.dim), dtype=torch.float64, device="cuda")
    dX = torch.zeros_like(dB)
    torch.cuda.synchronize()
    with pytest.raises(_lib.CosmicError, match="scratch"):
        M.solve_device_noalloc(1.0, dB.data_ptr(), M.dim, dX.data_ptr(), M.dim, r)
