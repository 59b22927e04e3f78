"""ssGBLUP H^-1 (src/io/relationship_builder.cpp:481-633, SplitMatrixAdapter
matrix_adapter.h:232-373): the device set-up against the oracle, the split
operator against the dense H^-1, and PCG/REML on it."""
import numpy as np
import pytest

from cosmicblup_amd import _lib, design, mme, ssgblup, synth
from cosmicblup_amd.pedigree import invertA_henderson
from oracle import grm as og
from oracle import reml as orl

pytestmark = pytest.mark.gpu


def problem(n_gen=6, per_gen=100, n_geno=200, m=1000, seed=7):
    dam, sire = synth.simulate_pedigree(n_gen, per_gen, seed=seed)
    n = dam.size
    gmap = np.arange(n - n_geno, n, dtype=np.int32)
    gmap[5] = -1  # a genotyped animal missing from the pedigree
    x, _ = synth.simulate_dosages(n_geno, m, maf=(0.05, 0.5), seed=seed + 1)
    G, _, _ = og.grm_vanraden(synth.encode_bed(x), n_geno, m)
    return dam, sire, gmap, G


def dense_hinv(ainv, gmap, D):
    H = ainv.toDense()
    for i, gi in enumerate(gmap):
        if gi < 0:
            continue
        for j, gj in enumerate(gmap):
            if gj >= 0:
                H[gi, gj] += D[i, j]
    return H


def test_blocks_match_oracle(cb, ctx):
    dam, sire, gmap, G = problem()
    D, A22, st = ssgblup.build_hinv_blocks(dam, sire, gmap, G, want_a22=True)
    Dr, A22r, (a, b) = orl.hinv_blocks(dam, sire, gmap, G)
    np.testing.assert_allclose(A22.cpu().numpy(), A22r, rtol=0, atol=1e-12)
    np.testing.assert_allclose([st["a"], st["b"]], [a, b], rtol=1e-12)
    Dn = D.cpu().numpy()
    assert np.abs(Dn - Dr).max() <= 1e-9 * np.abs(Dr).max()


def test_split_operator_solve_and_reml(cb, ctx):
    dam, sire, gmap, G = problem(seed=8)
    A = invertA_henderson(dam, sire)
    D, _, _ = ssgblup.build_hinv_blocks(dam, sire, gmap, G)
    n = dam.size
    rng = np.random.default_rng(3)
    rec = np.sort(rng.choice(n, 450, replace=False)).astype(np.int32)
    y = rng.standard_normal(rec.size) + 4.0
    fd = design.build_fixed_design([[f"g{int(v)}"] for v in rng.integers(0, 8, rec.size)])
    M = mme.MixedModel(fd, rec, y, n)
    M.set_qinv_split(A.ptr, A.col, A.val, gmap, D.data_ptr(), D.shape[0])
    H = dense_hinv(A, gmap, D.cpu().numpy())
    C, rhs, _ = orl.mme_dense(design.dense_x(fd), rec, y, H, 1.7)
    Xb = np.random.default_rng(0).standard_normal((M.dim, 33))
    for r in (1, 2, 33):
        Y = M.apply(1.7, Xb[:, :r])
        assert np.abs(Y - C @ Xb[:, :r]).max() <= 1e-12 * np.abs(C @ Xb[:, :r]).max()
    ref = np.linalg.solve(C, rhs)
    for prec in (_lib.CB_PREC_FP64, _lib.CB_PREC_MIXED, _lib.CB_PREC_MIXED_F16):
        X, _, _ = M.solve(1.7, rhs, mme.solve_options(precision=prec, tol=1e-12 if prec == 0 else 1e-10))
        assert np.abs(X[M.p:] - ref[M.p:]).max() <= 1e-6 * np.abs(ref[M.p:]).max()
    R = mme.AIREML(M, mme.reml_options(n_probes=16, seed=2, solve=mme.solve_options(tol=1e-10)))
    R.solve()
    o = orl.reml_run(design.dense_x(fd), rec, y, H, Zp=orl.probes(n, 16, 2))
    np.testing.assert_allclose([R.var_u, R.var_e], [o["vu"], o["ve"]], rtol=1e-4)
    np.testing.assert_allclose(R.logL(0.6, 0.9), orl.logL(design.dense_x(fd), rec, y, H, 0.6, 0.9), rtol=1e-10)
