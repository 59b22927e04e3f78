"""Phenotype / variance readers and the result files (.vars, .beta, .rand)
of a single-trait run: the reference's fixture (tests/golden/pblup_minimal,
copied from the reference's tests/fixtures), hand-checked formats and
closed forms. Host code only, except the end-to-end solve marked gpu."""
import math
from pathlib import Path

import numpy as np
import pytest

from cosmicblup_amd import design, phenotype, results
from cosmicblup_amd.pedigree import read_pedigree

GOLD = Path(__file__).parent / "golden" / "pblup_minimal"


def idmap_of(ids):
    return {x: i + 1 for i, x in enumerate(ids)}


def test_fixture_pheno_and_vars():
    ids, dam, sire = read_pedigree(str(GOLD / "pedigree.txt"))
    recs, st = phenotype.read_phen(str(GOLD / "pheno.txt"), idmap_of(ids), trait_name="trait",
                                   dcovar_names=["group"], qcovar_names=["cov"])
    assert [r.idstr for r in recs] == ["id1", "id2", "id3", "id4"]
    assert [r.aid for r in recs] == [1, 2, 3, 4]
    assert [r.y for r in recs] == [10.0, 12.0, 13.0, 14.0]
    assert [r.cats for r in recs] == [["A"], ["B"], ["A"], ["B"]]
    assert [r.nums for r in recs] == [[1.0], [2.0], [3.0], [4.0]]
    assert (st.n_rows_total, st.n_used, st.n_dropped_missing, st.n_not_in_id) == (4, 4, 0, 0)
    assert phenotype.read_vars(str(GOLD / "vars.txt")) == (2.0, 3.0)


def test_pheno_columns_delimiters_and_drops(tmp_path):
    # the reference test runner's toy phenotype: FID IID batch age Trait Trait2
    p = tmp_path / "toy.pheno"
    p.write_text("FID IID batch age Trait Trait2\n"
                 "0 sample1 A 10 10.0 20.1\n"
                 "0 sample2 A 11 NA 21.0\n"       # missing trait
                 "0 sample3 . 12 13.1 25.9\n"     # missing factor level
                 "0 sampleX B 13 14.0 28.2\n"     # not in the id map
                 "0 sample5 A 14 16.3abc\n"       # short row (Trait2 absent is fine, trait parses as 16.3)
                 "\n"
                 "0 sample6 B 15\n")              # short row
    idmap = idmap_of(["sample1", "sample2", "sample3", "sample5", "sample6"])
    dn, qn = [], []
    recs, st = phenotype.read_phen(str(p), idmap, trait_name="trait", dcovar_pos_1based=[3],
                                   qcovar_pos_1based=[4], dcovar_names=dn, qcovar_names=qn)
    assert dn == ["batch"] and qn == ["age"]
    assert [(r.idstr, r.y, r.cats, r.nums) for r in recs] == [("sample1", 10.0, ["A"], [10.0]),
                                                             ("sample5", 16.3, ["A"], [14.0])]
    assert (st.n_rows_total, st.n_dropped_missing, st.n_not_in_id, st.n_used) == (6, 3, 1, 2)
    # comma-delimited, id by explicit name, trait by 1-based position
    q = tmp_path / "c.csv"
    q.write_text("animal , y\n a1 , 1.5\na2,null\n")
    recs, st = phenotype.read_phen(str(q), {"a1": 1, "a2": 2}, trait_pos_1based=2, id_col_name="ANIMAL")
    assert [(r.idstr, r.y) for r in recs] == [("a1", 1.5)] and st.n_dropped_missing == 1
    with pytest.raises(RuntimeError, match="missing required ID column"):
        phenotype.read_phen(str(q), {}, trait_pos_1based=2)
    with pytest.raises(RuntimeError, match="Cannot find discrete covariate column: herd"):
        phenotype.read_phen(str(p), idmap, trait_name="Trait", dcovar_names=["herd"])


def test_read_vars_forms(tmp_path):
    p = tmp_path / "v.vars"
    p.write_text("Item\tVar\tVar_SE\nGA\t0.4\t0.1\nresidual\t0.6\t0.1\n")
    assert phenotype.read_vars(str(p)) == (0.4, 0.6)
    p.write_text("e 3\nu 2\n")
    assert phenotype.read_vars(str(p)) == (2.0, 3.0)
    p.write_text("x\n")
    with pytest.raises(RuntimeError, match="Cannot parse variance components"):
        phenotype.read_vars(str(p))


def test_vars_table(tmp_path):
    """h2 = vu / (vu + ve); its delta-method SE from the two SEs; the p value
    erfc(z / sqrt 2) printed as <2e-16 below that cutoff."""
    vu, ve, su, se = 0.4, 0.6, 0.05, 0.04
    path = results.write_vars(str(tmp_path / "run"), "gblup", vu, ve, su, se)
    lines = Path(path).read_text().splitlines()
    assert lines[0] == "Item\tVar\tVar_SE\th2\th2_SE\th2_Pr(Chisq)"
    tag, v, s, h2, sh2, pv = lines[1].split("\t")
    assert (tag, v, s) == ("GA", "0.4", "0.05")
    t = vu + ve
    want_se = math.sqrt(((t - vu) / t**2 * su) ** 2 + (vu / t**2 * se) ** 2)
    assert float(h2) == pytest.approx(0.4, rel=1e-6) and float(sh2) == pytest.approx(want_se, rel=1e-5)
    assert pv == "<2e-16"  # z = 0.4 / 0.0283 ~ 14
    e = lines[2].split("\t")
    assert e[0] == "e" and float(e[3]) == pytest.approx(0.6, rel=1e-6)
    assert results.format_pval(math.erfc(2.0 / math.sqrt(2))) == "0.0455003"
    assert results.calc_pval(1.0, 0.0) == 1.0
    assert Path(results.write_vars(str(tmp_path / "p"), "pblup", 1, 1)).read_text().splitlines()[1].startswith("PA\t")


def test_beta_and_rand_tables(tmp_path):
    """.beta: names, estimates, nan SE. .rand: animals with records in order
    of their first record, then the others by id; per animal the mean of u
    and of the residual over its records, and the count."""
    fd = design.build_fixed_design([["A"], ["B"], ["A"]], cat_names=["grp"])
    assert fd.names == ["mu", "grp_B"]
    beta = np.array([10.0, 2.0])
    bp = results.write_beta(str(tmp_path / "r"), fd, beta)
    assert Path(bp).read_text() == "Levels\tEstimation\tSE\nmu\t10\tnan\ngrp_B\t2\tnan\n"
    ids = ["z", "b", "a", "c"]
    u = np.array([0.5, -0.25, 0.125, 1.0])
    rec_aid = [2, 1, 2]            # records of b, z, b
    y = [12.0, 11.0, 10.5]
    rp = results.write_rand(str(tmp_path / "r"), "pblup", ids, rec_aid, y, fd, beta, u)
    rows = [l.split("\t") for l in Path(rp).read_text().splitlines()]
    assert rows[0] == ["ID", "PA", "residuals", "n_obs"]
    # b: records 0 and 2 (level A, x'b = 10): residuals 12 - (10 - 0.25) = 2.25
    # and 10.5 - (10 - 0.25) = 0.75, mean 1.5
    assert rows[1] == ["b", "-0.25", "1.5", "2"]
    # z: record 1 (B): 11 - 12 - 0.5
    assert rows[2] == ["z", "0.5", "-1.5", "1"]
    assert rows[3] == ["a", "0.125", "nan", "0"] and rows[4] == ["c", "1", "nan", "0"]


def test_obsrand_and_multi_trait_tables(tmp_path):
    """.obsrand: per record x'b, u of its level (0 without one), fitted and
    residual. .mv.vars: (i, j, Vg, Ve) lines. Multi-trait .beta / .rand: one
    column per trait, effect names of length p or Covar_k / Intercept."""
    fd = design.build_fixed_design([["A"], ["B"], ["A"]], cat_names=["grp"])
    beta = np.array([10.0, 2.0])
    u = np.array([0.5, -0.25])
    op = results.write_obsrand(str(tmp_path / "o"), "gblup", ["x1", "x2", "x3"], [11.0, 12.5, 9.0], fd, beta, u,
                               [2, 1, 0])
    rows = [l.split("\t") for l in Path(op).read_text().splitlines()]
    assert rows[0] == ["Record", "ID", "Y", "Fixed", "GA", "Fitted", "Residual"]
    assert rows[1] == ["1", "x1", "11", "10", "-0.25", "9.75", "1.25"]
    assert rows[2] == ["2", "x2", "12.5", "12", "0.5", "12.5", "0"]
    assert rows[3] == ["3", "x3", "9", "10", "0", "10", "-1"]
    vp = results.write_mv_vars(str(tmp_path / "m"), [[1.0, 0.25], [0.25, 2.0]], [[3.0, 0.5], [0.5, 4.0]])
    assert Path(vp).read_text().splitlines() == ["Trait1\tTrait2\tVg\tVe", "1\t1\t1\t3", "1\t2\t0.25\t0.5",
                                                 "2\t1\t0.25\t0.5", "2\t2\t2\t4"]
    bp, rp = results.write_multi_trait(str(tmp_path / "m"), [[1.5, -2.0], [0.125, 3.0]], [[0.1, 0.2], [0.3, 0.4]],
                                       ["a", "b"], trait_names=["ht", "wt"])
    assert Path(bp).read_text().splitlines() == ["Effect\tht\twt", "Covar_1\t1.5\t-2", "Covar_2\t0.125\t3"]
    assert Path(rp).read_text().splitlines() == ["ID\tht\twt", "a\t0.1\t0.2", "b\t0.3\t0.4"]
    bp, _ = results.write_multi_trait(str(tmp_path / "n"), [[7.0, 8.0]], [[0.0, 0.0]], ["a"],
                                      effect_names=["mu"])
    assert Path(bp).read_text().splitlines() == ["Effect\tTrait_1\tTrait_2", "mu\t7\t8"]


@pytest.mark.gpu
def test_pblup_minimal_end_to_end(cb, ctx, tmp_path):
    """The reference fixture's run (README: --model pblup with the given
    variances): pedigree, phenotype and variances read from its files,
    A^-1 and the MME solved on the device, .beta and .rand written; values
    against the dense oracle solve."""
    from cosmicblup_amd import mme
    from cosmicblup_amd.pedigree import invertA_henderson
    from oracle import reml as orl

    ids, dam, sire = read_pedigree(str(GOLD / "pedigree.txt"))
    recs, _ = phenotype.read_phen(str(GOLD / "pheno.txt"), idmap_of(ids), trait_name="trait")
    vu, ve = phenotype.read_vars(str(GOLD / "vars.txt"))
    fd = design.build_fixed_design([r.cats for r in recs])
    A = invertA_henderson(dam, sire)
    aid = np.array([r.aid - 1 for r in recs], dtype=np.int32)
    y = np.array([r.y for r in recs])
    M = mme.MixedModel(fd, aid, y, len(ids))
    M.set_qinv_csr(A.ptr, A.col, A.val)
    X, _, _ = M.solve(ve / vu, M.rhs(), mme.solve_options(tol=1e-12, precision=cb._lib.CB_PREC_FP64))
    C, rhs, _ = orl.mme_dense(design.dense_x(fd), aid, y, A.toDense(), ve / vu)
    ref = np.linalg.solve(C, rhs)
    np.testing.assert_allclose(X, ref, rtol=1e-10, atol=1e-12)
    # fixed-effect SEs (calc_se, on by default) by unit-column solves on the device
    _, se = M.beta_cov(ve / vu, ve, mme.solve_options(tol=1e-12, precision=cb._lib.CB_PREC_FP64))
    _, se_ref = orl.beta_cov(design.dense_x(fd), aid, y, A.toDense(), vu, ve)
    np.testing.assert_allclose(se, se_ref, rtol=1e-9)
    out = str(tmp_path / "pblup")
    bl = Path(results.write_beta(out, fd, X[:M.p], se)).read_text().splitlines()
    assert bl[0] == "Levels\tEstimation\tSE" and bl[1].split("\t")[0] == "mu"
    assert float(bl[1].split("\t")[1]) == pytest.approx(ref[0], rel=1e-5)
    assert float(bl[1].split("\t")[2]) == pytest.approx(se_ref[0], rel=1e-5)
    ol = Path(results.write_obsrand(out, "pblup", [r.idstr for r in recs], y, fd, X[:M.p], X[M.p:],
                                    aid + 1)).read_text().splitlines()
    assert ol[0] == "Record\tID\tY\tFixed\tPA\tFitted\tResidual" and len(ol) == 1 + len(recs)
    for k, line in enumerate(ol[1:]):
        rec, rid, yv, fx, g, fit, res = line.split("\t")
        assert int(rec) == k + 1 and rid == recs[k].idstr
        assert float(fit) == pytest.approx(float(fx) + float(g), rel=1e-5, abs=1e-9)
        assert float(res) == pytest.approx(y[k] - ref[0] - ref[M.p + aid[k]], rel=1e-5, abs=1e-9)
    rl = Path(results.write_rand(out, "pblup", ids, aid + 1, y, fd, X[:M.p], X[M.p:])).read_text().splitlines()
    assert [l.split("\t")[0] for l in rl[1:]] == ["id1", "id2", "id3", "id4"]
    for k, line in enumerate(rl[1:]):
        _, g, res, nobs = line.split("\t")
        assert float(g) == pytest.approx(ref[M.p + k], rel=1e-5, abs=1e-9) and nobs == "1"
        assert float(res) == pytest.approx(y[k] - ref[0] - ref[M.p + k], rel=1e-5, abs=1e-9)
