"""GPU parity of the GRM path (libcosmicblup_amd.so through the C-ABI) with
the oracle: bit-exact SNP statistics and integer Gram, FP64 G within 1e-12
of the FP64 oracle (max abs error; the reference's own FP64 GEMM rounding
is of the same order), FP32 G equal to the rounded FP64 G."""
import numpy as np
import pytest

from cosmicblup_amd import synth
from oracle import grm as og

pytestmark = pytest.mark.gpu

TOL_G64 = 1e-12  # max |G_gpu - G_oracle|


def run_gpu(cb, bed, n, m, outputs=None, **opts):
    outputs = outputs or (cb.CB_GRM_OUT_F64 | cb.CB_GRM_OUT_F32 | cb.CB_GRM_OUT_INT)
    r = cb.PlinkReader()
    r.from_bytes(bed, n, m)
    gb = cb.GrmBuilder(cb.Options(**opts))
    gb.compute(r, outputs=outputs)
    return gb


def check_case(cb, x, block_size=2048, min_maf=0.01):
    n, m = x.shape
    bed = synth.encode_bed(x)
    gb = run_gpu(cb, bed, n, m, block_size=block_size, min_maf=min_maf)
    st = og.snp_stats(bed, n, m, min_maf=min_maf)
    gs = gb.getSnpStats()
    for k in ("n_obs", "sum_x", "n_hom1", "n_het", "n_hom2", "valid"):
        assert np.array_equal(gs[k].astype(np.int64), st[k].astype(np.int64)), k
    assert np.array_equal(gs["mean_f"], st["mean_f"])
    assert np.array_equal(gs["twopq_f"], st["twopq_f"])
    assert np.array_equal(gb.getAlleleFreqs(), st["freq_f"])
    Gi = gb.getIntGram().astype(np.int64)
    assert np.array_equal(Gi, og.int_gram(bed, n, m, st["valid"]))
    G_ref, S_ref, nv_ref = og.grm_vanraden(bed, n, m, block_size=block_size, min_maf=min_maf)
    assert gb.getSum2pq() == S_ref
    assert gb.getTotalSnps() == nv_ref
    G = gb.getGrm()
    err = np.abs(G - G_ref).max()
    assert err < TOL_G64, err
    assert np.array_equal(gb.getGrmF32(), G.astype(np.float32))
    assert np.array_equal(G, G.T)
    return err


def test_toy_fixture(cb, ctx):
    from tests.test_oracle_grm import TOY_DOSAGE

    check_case(cb, TOY_DOSAGE)


@pytest.mark.parametrize("n,m,miss,seed", [
    (256, 256, 0.0, 1),      # exactly one tile / one stage
    (300, 700, 0.0, 2),      # ragged n and m
    (517, 1029, 0.01, 3),    # ragged, missing calls
    (1000, 3000, 0.05, 4),   # heavier missingness, several tiles
    (5, 17, 0.1, 5),         # tiny
    # the Gram epilogue moves 2 entries per lane when n is even, else 1:
    # ragged even n with and without missing calls (517 and 5 take the
    # 1-entry path)
    (770, 1500, 0.02, 6),
    (302, 500, 0.0, 7),
])
def test_random_cases(cb, ctx, n, m, miss, seed):
    x, _ = synth.simulate_dosages(n, m, maf=(0.0, 0.5), missing_rate=miss, seed=seed)
    check_case(cb, x, block_size=512)


def test_long_missing_lists(cb, ctx):
    """Samples 0 and 3 missing at all 300,000 SNPs, sample 5 at every other
    one. Their lists outrun one fold of the missing-call kernel's digit sums
    (32,640 entries of a sample), so the kernel must add its partial sums
    into E and start again, each fold with its own sum of K for sample 3's
    code offset; the pair (0, 3) shares every missing SNP, the largest
    both-missing sum."""
    x, _ = synth.simulate_dosages(64, 300_000, maf=(0.05, 0.5), seed=17)
    x[0, :] = -1
    x[3, :] = -1
    x[5, ::2] = -1
    check_case(cb, x, block_size=4096)


def test_config1_shape(cb, ctx):
    # n=2000, m=10000, MAF U(0.05,0.5), no missing calls
    x, _ = synth.simulate_dosages(2000, 10000, maf=(0.05, 0.5), seed=20261203)
    check_case(cb, x)


def test_maf_boundary_and_dropped_snps(cb, ctx):
    """SNPs whose p straddles the 0.01f MAF literal from both sides (minor
    and major allele counted), an all-missing SNP and a monomorphic one,
    among random SNPs with missing calls: the device keeps exactly the SNPs
    the oracle keeps, and G matches."""
    n = 1000
    rng = np.random.default_rng(17)
    sums = [19, 20, 21, 1979, 1980, 1981, 0, 2000]
    xb = np.zeros((n, len(sums)), dtype=np.int8)
    for j, sm in enumerate(sums):
        xb[: sm // 2, j] = 2
        if sm % 2:
            xb[sm // 2, j] = 1
        rng.shuffle(xb[:, j])
    x_rand, _ = synth.simulate_dosages(n, 120, maf=(0.02, 0.5), missing_rate=0.02, seed=5)
    x_miss = np.full((n, 1), -1, dtype=np.int8)
    x = np.concatenate([xb[:, :3], x_rand[:, :60], x_miss, xb[:, 3:], x_rand[:, 60:]], axis=1)
    check_case(cb, x)
    st = og.snp_stats(synth.encode_bed(x), n, x.shape[1])
    kept = st["valid"].astype(bool)
    # p = 19/2000 drops, 20/2000 (exactly 0.01 > (double)0.01f) and 21/2000
    # stay; the same counted as the major allele (1979, 1980, 1981 of 2000);
    # monomorphic 0 and 2000; the all-missing SNP drops
    assert list(kept[:3]) == [False, True, True]
    assert not kept[63]
    assert list(kept[64:69]) == [True, True, False, False, False]


def test_all_snps_filtered(cb, ctx):
    x = np.full((64, 40), 2, dtype=np.int8)  # monomorphic everywhere
    bed = synth.encode_bed(x)
    gb = run_gpu(cb, bed, 64, 40)
    assert gb.getTotalSnps() == 0 and gb.getSum2pq() == 0.0
    assert not gb.getGrm().any()


def test_rejects_bad_inputs(cb, ctx):
    x, _ = synth.simulate_dosages(10, 10, seed=1)
    bed = synth.encode_bed(x)
    bad = bed.copy()
    bad[0] = 0
    r = cb.PlinkReader()
    with pytest.raises(cb.CosmicError, match="magic"):
        r.from_bytes(bad, 10, 10)
    with pytest.raises(cb.CosmicError, match="size"):
        r.from_bytes(bed[:-1], 10, 10)
    with pytest.raises(cb.CosmicError, match="min_maf"):
        cb.GrmBuilder(cb.Options(min_maf=0.001))


def test_large_properties(cb, ctx):
    """n=8192, m=20000 with 1% missing: symmetry, exact integer diagonal,
    and rows spot-checked against the oracle's FP64 block."""
    n, m = 8192, 20000
    x, _ = synth.simulate_dosages(n, m, maf=(0.01, 0.5), missing_rate=0.01, seed=9)
    bed = synth.encode_bed(x)
    gb = run_gpu(cb, bed, n, m, outputs=cb.CB_GRM_OUT_F64)
    G = gb.getGrm()
    assert np.array_equal(G, G.T)
    st = og.snp_stats(bed, n, m)
    rows = [0, 1, 4095, 8191]
    S = gb.getSum2pq()
    for r0 in rows:
        blk = og.grm_block_raw(bed, n, m, st, r0, 1, 0, n) / S
        assert np.abs(G[r0] - blk[0]).max() < TOL_G64


def test_grm_written_from_device(cb, ctx, tmp_path):
    """cb_grm_write streams the FP64 G from the device in row panels (n =
    3000 spans three): the COSMIC_UPPER_PACKED file holds exactly getGrm(),
    and the text file holds its lower triangle in the stream format."""
    from cosmicblup_amd import matrixio

    n, m = 3000, 2000
    x, _ = synth.simulate_dosages(n, m, maf=(0.05, 0.5), missing_rate=0.01, seed=21)
    gb = run_gpu(cb, synth.encode_bed(x), n, m, outputs=cb.CB_GRM_OUT_F64)
    G = gb.getGrm()
    pb = tmp_path / "g.GA.bin"
    gb.writeMatrix(str(pb), "bin")
    assert pb.stat().st_size == 32 + 28 + 8 * n * (n + 1) // 2
    np.testing.assert_array_equal(matrixio.read_upper_packed_binary(str(pb)), G)
    # text: spot-check rows from every panel against the stream format
    pt = tmp_path / "g.GA.txt"
    gb.writeMatrix(str(pt), "txt")
    with open(pt) as f:
        lines = f.read().splitlines()
    assert len(lines) == n * (n + 1) // 2
    for i in (0, 1, 1397, 1398, 2796, 2999):
        start = i * (i + 1) // 2
        assert lines[start:start + i + 1] == [f"{i + 1}\t{j + 1}\t{G[i, j]:g}" for j in range(i + 1)]
