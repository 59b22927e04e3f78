"""Pins the GRM oracle (oracle/grm_oracle.c, oracle/grm.py) against facts that
do not restate it: the hand-decoded toy fixture of the reference's tests,
closed-form counts, brute-force integer products, the MAF threshold in float,
and the G 1 invariant that separates float- from double-rounded centring."""
import numpy as np
import pytest

from cosmicblup_amd import synth
from oracle import grm as og

ROOT = __import__("pathlib").Path(__file__).resolve().parent
TOY = ROOT / "golden" / "toy_plink"

# Decoded by hand from tests/fixtures/toy_plink/toy.bed of the reference
# (bytes cb 02 | 3a 08 | e0 0e | 8b 03, sample i at bits 2(i&3)).
TOY_DOSAGE = np.array([[0, 1, 2, 0, 1, 2],
                       [1, 1, 0, 2, 2, 1],
                       [2, 2, 1, 0, 1, 0],
                       [0, 1, 2, 1, 0, 2]], dtype=np.int8).T


def toy_bed():
    return np.fromfile(TOY / "toy.bed", dtype=np.uint8)


def test_toy_fixture_decodes_to_known_dosages():
    bed = toy_bed()
    assert bed.size == 3 + 4 * 2
    assert np.array_equal(synth.decode_bed(bed, 6, 4), TOY_DOSAGE)
    # writer round trip reproduces the reference's file byte for byte
    assert np.array_equal(synth.encode_bed(TOY_DOSAGE), bed)


def test_toy_stats_and_grm():
    bed = toy_bed()
    st = og.snp_stats(bed, 6, 4)
    sums = TOY_DOSAGE.sum(axis=0)
    assert np.array_equal(st["sum_x"], sums)
    assert np.array_equal(st["n_obs"], [6, 6, 6, 6])
    assert np.array_equal(st["n_hom1"], (TOY_DOSAGE == 2).sum(0))
    assert np.array_equal(st["n_het"], (TOY_DOSAGE == 1).sum(0))
    p = sums / 12.0
    assert np.array_equal(st["freq_f"], p.astype(np.float32))
    G, S, nv = og.grm_vanraden(bed, 6, 4)
    assert nv == 4
    assert S == float(np.sum((2 * p * (1 - p)).astype(np.float32).astype(np.float64)))
    mu = (2 * p).astype(np.float32).astype(np.float64)
    Z = TOY_DOSAGE - mu
    np.testing.assert_allclose(G, Z @ Z.T / S, rtol=0, atol=1e-14)


@pytest.mark.parametrize("n,m,miss", [(37, 50, 0.0), (64, 33, 0.05), (129, 300, 0.02)])
def test_stats_closed_form(n, m, miss):
    x, _ = synth.simulate_dosages(n, m, maf=(0.0, 0.5), missing_rate=miss, seed=n + m)
    bed = synth.encode_bed(x)
    st = og.snp_stats(bed, n, m)
    obs = (x >= 0).sum(0)
    assert np.array_equal(st["n_obs"], obs)
    assert np.array_equal(st["sum_x"], np.where(x > 0, x, 0).sum(0))
    assert np.array_equal(st["n_hom2"], (x == 0).sum(0))
    assert np.array_equal(st["n_hom1"] + st["n_het"] + st["n_hom2"], obs)


def test_maf_threshold_is_the_float_literal():
    # n_obs = 1000 per SNP; sum_x chosen so p straddles 0.01 and (double)0.01f
    n = 1000
    sums = [19, 20, 21, 0, 2000, 1980, 1981]
    m = len(sums)
    x = np.zeros((n, m), dtype=np.int8)
    for j, s in enumerate(sums):
        x[: s // 2, j] = 2
        if s % 2:
            x[s // 2, j] = 1
    st = og.snp_stats(synth.encode_bed(x), n, m)
    p = np.array(sums) / 2000.0
    maf = np.minimum(p, 1 - p)
    expect = (maf >= np.float64(np.float32(0.01))) & (np.abs(maf) >= 1e-8)
    assert np.array_equal(st["valid"].astype(bool), expect)
    # p = 0.01 exactly passes because 0.01 > (double)0.01f
    assert st["valid"][1] == 1 and st["valid"][0] == 0
    assert np.all(st["freq_f"][~expect] == -1.0)


def test_all_missing_and_monomorphic_snps_are_dropped():
    n, m = 40, 3
    x = np.ones((n, m), dtype=np.int8)
    x[:, 0] = -1
    x[:, 1] = 2
    st = og.snp_stats(synth.encode_bed(x), n, m)
    assert list(st["valid"]) == [0, 0, 1]
    assert st["n_obs"][0] == 0


@pytest.mark.parametrize("miss", [0.0, 0.03])
def test_int_gram_brute_force(miss):
    n, m = 70, 90
    x, _ = synth.simulate_dosages(n, m, maf=(0.0, 0.5), missing_rate=miss, seed=3)
    bed = synth.encode_bed(x)
    st = og.snp_stats(bed, n, m)
    X = np.where(x > 0, x, 0).astype(np.int64) * st["valid"].astype(np.int64)
    assert np.array_equal(og.int_gram(bed, n, m, st["valid"]), X @ X.T)


def test_grm_centring_uses_float_rounded_mean():
    # Without missing calls sum_i Z_ij = n (2p_j - mu_j); with mu_j the float
    # rounding of 2p_j this is ~1e-8, so G 1 is predictably non-zero.
    n, m = 200, 400
    x, _ = synth.simulate_dosages(n, m, maf=(0.05, 0.5), seed=11)
    bed = synth.encode_bed(x)
    G, S, nv = og.grm_vanraden(bed, n, m, block_size=64)
    st = og.snp_stats(bed, n, m)
    v = st["valid"].astype(bool)
    p = st["sum_x"][v] / (2.0 * st["n_obs"][v])
    mu_f = (2 * p).astype(np.float32).astype(np.float64)
    Zf = x[:, v] - mu_f
    pred = Zf @ (n * (2 * p - mu_f)) / S
    np.testing.assert_allclose(G.sum(1), pred, rtol=0, atol=1e-11)
    # and the double-rounded alternative would be measurably different
    assert np.abs(pred).max() > 1e-9


def test_grm_matches_matrix_product_with_missing():
    n, m = 150, 500
    x, _ = synth.simulate_dosages(n, m, maf=(0.0, 0.5), missing_rate=0.02, seed=5)
    bed = synth.encode_bed(x)
    G, S, nv = og.grm_vanraden(bed, n, m, block_size=128)
    st = og.snp_stats(bed, n, m)
    v = st["valid"].astype(bool)
    mu = st["mean_f"][v].astype(np.float64)
    xv = x[:, v].astype(np.float64)
    Z = np.where(xv >= 0, xv - mu, 0.0)  # missing calls centred to 0
    assert nv == v.sum()
    # S: double sum of float 2pq in blocks of 128 SNPs
    tw = st["twopq_f"].astype(np.float64) * v
    S_ref = sum(float(np.sum(tw[b:b + 128])) for b in range(0, m, 128))
    assert abs(S - S_ref) <= 1e-12 * S
    np.testing.assert_allclose(G, Z @ Z.T / S, rtol=0, atol=1e-13)


def test_invert_psd_identity_and_floor():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((30, 30))
    M = A @ A.T + 30 * np.eye(30)
    np.testing.assert_allclose(og.invert_psd_matrix(M) @ M, np.eye(30), atol=1e-12)
    # centred G is singular along 1: that eigenvalue is floored at 1e-6
    x, _ = synth.simulate_dosages(40, 200, seed=2)
    bed = synth.encode_bed(x)
    G, _, _ = og.grm_vanraden(bed, 40, 200)
    Gi = og.invert_psd_matrix(G)
    one = np.ones(40) / np.sqrt(40)
    assert one @ Gi @ one > 1e5
    Gb = og.bent_matrix(G)
    np.testing.assert_allclose(Gb @ Gi, np.eye(40), atol=1e-6)


@pytest.mark.parametrize("miss", [0.0, 0.03])
def test_block_and_row_oracles_match_the_full_grm(miss):
    """The sub-block and row-list forms (used for single rows of a G too
    large for the full oracle, and as the CPU baseline) against the full
    orc_grm_vanraden times S and a brute-force numpy Z Z^T, including a
    partial last SNP block and non-contiguous rows."""
    n, m = 150, 4500
    x, _ = synth.simulate_dosages(n, m, maf=(0.01, 0.5), missing_rate=miss, seed=31)
    bed = synth.encode_bed(x)
    st = og.snp_stats(bed, n, m)
    G, S, _ = og.grm_vanraden(bed, n, m)
    v = st["valid"].astype(bool)
    mu = st["mean_f"].astype(np.float64)
    Z = np.where(x >= 0, x - mu, 0.0)[:, v]
    ZZ = Z @ Z.T
    blk = og.grm_block_raw(bed, n, m, st, 17, 40, 60, 90)
    np.testing.assert_allclose(blk, ZZ[17:57, 60:150], rtol=0, atol=1e-9)
    np.testing.assert_allclose(blk / S, G[17:57, 60:150], rtol=0, atol=1e-14)
    sel = np.array([0, 1, 77, 149, 3])
    rr = og.grm_rows_raw(bed[3:].reshape(m, -1), (n + 3) // 4, n, m, st, sel)
    assert np.array_equal(rr, og.grm_block_raw(bed, n, m, st, 0, n, 0, n)[sel])
    np.testing.assert_allclose(rr, ZZ[sel], rtol=0, atol=1e-9)
    # rows padded to a wider pitch (device row layout) give the same sums
    rows = np.zeros((m, 48), dtype=np.uint8)
    rows[:, :(n + 3) // 4] = bed[3:].reshape(m, -1)
    st2 = og.snp_stats_rows(rows, 48, n, m)
    assert all(np.array_equal(st[k], st2[k]) for k in st)
    assert np.array_equal(og.grm_rows_raw(rows, 48, n, m, st2, sel), rr)
