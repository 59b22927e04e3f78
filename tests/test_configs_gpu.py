"""The configurations at their real sizes, checked against the oracle.

  GRM at scale: n = m = 50,000, MAF U(0.01, 0.5), 1% MCAR missing calls.
  Index ranges past n^2 = 2^31 (n >= 46,341) are reached only at this size.
  The device G is checked for exact symmetry, its integer Gram diagonal
  against sum_j x_ij^2 over the kept SNPs (exact), SNP statistics bit for
  bit, and rows straddling 2^31 against the oracle's FP64 Z Z^T / sum(2pq).

  Plain GBLUP with dense-exact AI-REML: n = 2,000, m = 10,000, MAF
  U(0.05, 0.5), 20 contemporary groups, h2 = 0.3, from (1, 1) to
  sum |d sigma^2| < 1e-4, against the oracle's run on the same G^-1.
"""
import ctypes as C

import numpy as np
import pytest

from cosmicblup_amd import _lib, design, mme, synth
from oracle import grm as og
from oracle import reml as orl

pytestmark = pytest.mark.gpu

TOL_G64 = 1e-12


def _hip():
    return C.CDLL("libamdhip64.so")


def _device_diag_i32(ptr: int, n: int, ld: int) -> np.ndarray:
    """Diagonal of an n x n int32 device matrix (leading dimension ld)."""
    out = np.empty(n, dtype=np.int32)
    hip = _hip()
    hip.hipMemcpy2D.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_int]
    rc = hip.hipMemcpy2D(out.ctypes.data, 4, ptr, (ld + 1) * 4, 4, n, 2)  # hipMemcpyDeviceToHost
    assert rc == 0, f"hipMemcpy2D failed ({rc})"
    return out


@pytest.mark.timeout(900)
def test_grm_n50000_m50000_missing_1pct(cb):
    import torch

    n = m = 50000
    print(f"[grm-full] generating BED n={n} m={m}", flush=True)
    rows, pitch = synth.bed_on_device(n, m, maf=(0.01, 0.5), missing_rate=0.01, seed=20262203)
    ctx = _lib.default_context()
    reader = cb.PlinkReader.from_device_rows(ctx, rows.data_ptr(), pitch, n, m)
    gb = cb.GrmBuilder(cb.Options(), ctx=ctx)
    gb.compute(reader, outputs=cb.CB_GRM_OUT_F64 | cb.CB_GRM_OUT_INT)
    print("[grm-full] computed", flush=True)
    host_rows = rows.cpu().numpy()
    st = og.snp_stats_rows(host_rows, pitch, n, m)
    gs = gb.getSnpStats()
    for k_dev, k_ref in (("n_obs", "n_obs"), ("sum_x", "sum_x"), ("n_hom1", "n_hom1"), ("n_het", "n_het"),
                         ("n_hom2", "n_hom2"), ("valid", "valid"), ("mean_f", "mean_f"), ("twopq_f", "twopq_f")):
        assert np.array_equal(gs[k_dev], st[k_ref]), k_dev
    valid = torch.from_numpy(st["valid"].astype(bool)).cuda()
    # integer Gram diagonal = sum over kept SNPs of x^2 (codes 00 -> 4, 10 -> 1, else 0)
    sq = torch.tensor([4, 0, 1, 0], dtype=torch.int32, device="cuda")
    kept = rows[valid]
    diag = torch.zeros(4 * pitch, dtype=torch.int64, device="cuda")
    for t in range(4):
        diag[t::4] = sq[((kept >> (2 * t)) & 3).long()].sum(0, dtype=torch.int64)
    del kept
    iptr, ild = gb.devicePtr(cb.CB_GRM_OUT_INT)
    torch.cuda.synchronize()
    assert np.array_equal(_device_diag_i32(iptr, n, ild).astype(np.int64), diag[:n].cpu().numpy())
    # FP64 G: exact symmetry in row panels, then rows against the oracle
    G = torch.empty((n, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    gb.copyGrmToDevice(G.data_ptr(), n)
    for i0 in range(0, n, 5000):
        assert torch.equal(G[i0:i0 + 5000], G[:, i0:i0 + 5000].T), i0
    print("[grm-full] symmetric", flush=True)
    sel = np.array([0, 1, 25000, 46340, 46341, 49999])
    got = G[torch.from_numpy(sel).cuda()].cpu().numpy()
    del G
    # sum(2pq): double sums of float 2pq per 2048-SNP block, then across blocks
    S_ref = 0.0
    for b0 in range(0, m, 2048):
        blk = 0.0
        for j in range(b0, min(m, b0 + 2048)):
            if st["valid"][j]:
                blk += float(st["twopq_f"][j])
        S_ref += blk
    S = gb.getSum2pq()
    assert S == S_ref
    ref = og.grm_rows_raw(host_rows, pitch, n, m, st, sel) / S
    err = np.abs(got - ref).max()
    print(f"[grm-full] rows {list(sel)}: max |dG| = {err:.3e}", flush=True)
    assert err < TOL_G64


def c1_problem(n=2000, m=10000, ngroups=20, h2=0.3, seed=20261203):
    x, _ = synth.simulate_dosages(n, m, maf=(0.05, 0.5), seed=seed)
    bed = synth.encode_bed(x)
    G, _, _ = og.grm_vanraden(bed, n, m)
    Q = og.invert_psd_matrix(G)
    y, cg = synth.simulate_phenotype(x, n_qtl=500, h2=h2, n_groups=ngroups, seed=seed + 1)
    fd = design.build_fixed_design([[c] for c in cg])
    return fd, np.arange(n, dtype=np.int32), y, Q


@pytest.mark.timeout(900)
def test_c1_dense_exact_reml_n2000_matches_oracle(cb, ctx):
    fd, aid, y, Q = c1_problem()
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    R = mme.AIREML(M, mme.reml_options(algorithm=_lib.CB_VCE_AI_EXACT, max_iter=20, tol=1e-4))
    R.solve()
    o = orl.dense_exact_run(design.dense_x(fd), aid, y, Q, max_iter=20, tol=1e-4)
    print(f"[c1] device {R.iterations} iterations vu={R.var_u:.8f} ve={R.var_e:.8f}; "
          f"oracle {len(o['hist'])} vu={o['vu']:.8f} ve={o['ve']:.8f}", flush=True)
    assert R.converged and o["converged"] and R.iterations == len(o["hist"])
    assert abs(R.var_u - o["vu"]) / o["vu"] < 1e-6 and abs(R.var_e - o["ve"]) / o["ve"] < 1e-6
    u, ur = R.solution[M.p:], o["sol"][M.p:]
    assert np.abs(u - ur).max() / np.abs(ur).max() < 1e-6


C3 = dict(seed=20263203, h2=0.25, n_rec=600_000, n_groups=2000, n_geno=20_000, m_geno=40_000)


@pytest.fixture(scope="module")
def c3_problem(cb):
    """The 1,000,000-animal pedigree MME (10 generations of 100,000, parents
    uniform from the previous generation, 5% unknown; A^-1 by Henderson's
    rules with inbreeding; 600,000 records in 2000 contemporary groups; h2 =
    0.25) and its W'W + lambda [0 0; 0 A^-1] and W'y assembled independently
    on the host with scipy (W = [X Z])."""
    import scipy.sparse as sp

    from cosmicblup_amd.pedigree import invertA_henderson

    seed, h2, n_rec = C3["seed"], C3["h2"], C3["n_rec"]
    dam, sire = synth.simulate_pedigree(10, 100_000, unknown=0.05, seed=seed)
    q = dam.size
    A = invertA_henderson(dam, sire)
    rec, y, grp = synth.simulate_pedigree_phenotype(dam, sire, A.dii, n_rec, h2=h2, n_groups=C3["n_groups"],
                                                    seed=seed + 1)
    fd = design.build_fixed_design([[f"cg{int(c):04d}"] for c in grp])
    lam = (1.0 - h2) / h2
    X = sp.csr_matrix((fd.val, fd.col, fd.ptr), shape=(n_rec, fd.p))
    Z = sp.csr_matrix((np.ones(n_rec), np.asarray(rec, dtype=np.int64), np.arange(n_rec + 1)), shape=(n_rec, q))
    W = sp.hstack([X, Z]).tocsr()
    Ai = sp.csr_matrix((A.val, A.col, A.ptr), shape=(q, q))
    Cm = (W.T @ W + sp.block_diag([sp.csr_matrix((fd.p, fd.p)), lam * Ai])).tocsr()
    b = W.T @ np.asarray(y, dtype=np.float64)
    return dict(dam=dam, sire=sire, A=A, rec=rec, y=y, fd=fd, lam=lam, q=q, Cm=Cm, b=b)


def _c3_solves(model, lam, b, C_apply, p):
    sols = {}
    for name, prec in (("fp64", _lib.CB_PREC_FP64), ("mixed", _lib.CB_PREC_MIXED)):
        x, it, rel = model.solve(lam, b, mme.solve_options(precision=prec, tol=1e-12, max_iter=20000))
        res = np.linalg.norm(C_apply(x) - b) / np.linalg.norm(b)
        print(f"[c3] {name}: {it} iterations, device rel {rel:.2e}, host rel {res:.2e}", flush=True)
        assert res < 1e-11, (name, res)
        sols[name] = x
    scale = np.abs(sols["fp64"][p:]).max()
    assert np.abs(sols["mixed"][p:] - sols["fp64"][p:]).max() / scale < 1e-6


@pytest.mark.timeout(900)
def test_c3_pblup_1m_animals_solution_satisfies_independent_mme(cb, ctx, c3_problem):
    """PBLUP at C3's size: the device solutions (FP64, and the mixed FP32
    recursion with FP64 residual replacement, both to relative residual
    1e-12) have relative residual ||C x - b|| / ||b|| below 1e-11 in the MME
    assembled on the host, and the two precisions' EBVs agree to 1e-6."""
    P = c3_problem
    model = mme.MixedModel(P["fd"], P["rec"], P["y"], P["q"], ctx=ctx)
    A = P["A"]
    model.set_qinv_csr(A.ptr, A.col, A.val)
    b = P["b"]
    b_dev = model.rhs()
    np.testing.assert_allclose(b_dev, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())
    _c3_solves(model, P["lam"], b, lambda x: P["Cm"] @ x, P["fd"].p)


@pytest.mark.timeout(900)
def test_c3_ssgblup_20k_genotyped_solution_satisfies_independent_mme(cb, ctx, c3_problem):
    """ssGBLUP at C3's size: the last 20,000 animals genotyped at 40,000 SNPs
    (GRM on the device), H^-1 = A^-1 + [0 0; 0 D] with D = G^-1 - A22^-1 on
    the genotyped rows (tuneG and blending on the device). The device
    solutions with the split operator (CSR A^-1 plus the gathered dense D)
    satisfy C = W'W + lambda H^-1, applied on the host as the scipy MME plus
    lambda P' D P x with D copied from the device, to 1e-11."""
    import torch

    from cosmicblup_amd import grm as cgrm
    from cosmicblup_amd import plink, ssgblup

    P = c3_problem
    q, p, lam = P["q"], P["fd"].p, P["lam"]
    ng, mg = C3["n_geno"], C3["m_geno"]
    gmap = np.arange(q - ng, q, dtype=np.int32)
    rows, pitch = synth.bed_on_device(ng, mg, maf=(0.01, 0.5), missing_rate=0.0, seed=C3["seed"] + 2)
    reader = plink.PlinkReader.from_device_rows(ctx, rows.data_ptr(), pitch, ng, mg)
    gb = cgrm.GrmBuilder(cgrm.Options(), ctx=ctx)
    gb.compute(reader, outputs=_lib.CB_GRM_OUT_F64)
    G = torch.empty((ng, ng), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    gb.copyGrmToDevice(G.data_ptr(), ng)
    del gb, reader, rows
    D, _, _ = ssgblup.build_hinv_blocks(P["dam"], P["sire"], gmap, G, ctx=ctx)
    del G
    model = mme.MixedModel(P["fd"], P["rec"], P["y"], q, ctx=ctx)
    A = P["A"]
    model.set_qinv_split(A.ptr, A.col, A.val, gmap, D.data_ptr(), ng)
    Dh = D.cpu().numpy()  # row-major ng x ng
    del D
    torch.cuda.empty_cache()
    g_rows = p + gmap

    def C_apply(x):
        y = P["Cm"] @ x
        y[g_rows] += lam * (Dh @ x[g_rows])
        return y

    _c3_solves(model, lam, P["b"], C_apply, p)
