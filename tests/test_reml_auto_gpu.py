"""The reference's size-based VCE branch (CB_VCE_AUTO) and the REML score
at the n = 20,000 shape.

Branch table (AI_REML::run_ai_iteration, vce.cpp:1334-1574; build_lhs's
solvers, vce.cpp:1041-1110; Ainv factorisation, vce.cpp:877-1001;
estimate_trace_Qinv_Cuu, vce.cpp:1266-1315; mc_samples 100):
  dim < 5000                                  dense-exact every iteration
  dense Q, 5000 <= dim < 15000, q > 5000      EM, then AI with logL check; adaptive 33..100 probes
  dense Q, 5000 <= dim < 15000, q <= 5000     the same with exact traces
  dense Q, dim >= 15000                       EM; adaptive 30..100 probes
  CSR Q, q <= 100000, dim < 200000            AI from the start with logL check; adaptive 33..100
  CSR Q, q > 100000, dim < 200000             EM, then AI (sparse direct solver) without logL here
  CSR Q, dim >= 200000                        EM; adaptive 30..100
"""
import numpy as np
import pytest

from cosmicblup_amd import _lib, design, mme, synth
from oracle import grm as og
from oracle import reml as orl

pytestmark = pytest.mark.gpu

AI, EM, EXACT = _lib.CB_VCE_AI, _lib.CB_VCE_EM, _lib.CB_VCE_AI_EXACT


def _model(n, q, ngroups, qkind, seed=1):
    import torch

    rng = np.random.default_rng(seed)
    grp = rng.integers(0, ngroups, n)
    width = len(str(ngroups))
    fd = design.build_fixed_design([[f"g{int(c):0{width}d}"] for c in grp])
    aid = (np.arange(n) % q).astype(np.int32)
    M = mme.MixedModel(fd, aid, rng.standard_normal(n), q)
    if qkind == "dense":
        Q = torch.eye(q, dtype=torch.float64, device="cuda") * 2.0
        torch.cuda.synchronize()
        M.set_qinv_dense_device(Q.data_ptr(), q)
    else:
        M.set_qinv_csr(np.arange(q + 1, dtype=np.int64), np.arange(q, dtype=np.int32), np.full(q, 2.0))
    return M


@pytest.mark.parametrize("n,q,ngroups,qkind,mc,want", [
    (3000, 3000, 10, "dense", 100, (EXACT, EXACT, 2, True, 0, 0)),
    (6000, 6000, 10, "dense", 100, (EM, AI, 1, True, 33, 100)),
    (4500, 4500, 700, "dense", 100, (EM, AI, 2, True, 0, 0)),
    (16000, 16000, 10, "dense", 100, (EM, EM, 1, False, 30, 100)),
    (16000, 16000, 10, "dense", 300, (EM, EM, 1, False, 50, 150)),
    (8000, 8000, 10, "csr", 100, (AI, AI, 1, True, 33, 100)),
    (120000, 120000, 10, "csr", 100, (EM, AI, 1, False, 33, 100)),
    (250000, 250000, 10, "csr", 100, (EM, EM, 1, False, 30, 100)),
])
def test_auto_branch_table(cb, ctx, n, q, ngroups, qkind, mc, want):
    M = _model(n, q, ngroups, qkind)
    b = mme.AIREML(M, mme.reml_options(algorithm=_lib.CB_VCE_AUTO, mc_samples=mc)).branch()
    got = (b["first"], b["later"], b["trace"], b["logl_checked"], b["min_probes"], b["max_probes"])
    assert got == want, (M.dim, got)


def test_explicit_algorithms_keep_their_steps(cb, ctx):
    M = _model(6000, 6000, 10, "dense")
    b = mme.AIREML(M, mme.reml_options(algorithm=EM, adaptive=True)).branch()
    assert (b["first"], b["later"], b["trace"], b["min_probes"], b["max_probes"]) == (EM, EM, 1, 30, 100)
    b = mme.AIREML(M, mme.reml_options(algorithm=AI, n_probes=64)).branch()
    assert (b["first"], b["later"], b["trace"], b["logl_checked"]) == (AI, AI, 0, False)


def gblup(n, m, ngroups=10, h2=0.35, seed=3):
    x, _ = synth.simulate_dosages(n, m, maf=(0.05, 0.5), seed=seed)
    G, _, _ = og.grm_vanraden(synth.encode_bed(x), n, m)
    Q = og.invert_psd_matrix(G)
    y, cg = synth.simulate_phenotype(x, n_qtl=60, h2=h2, n_groups=ngroups, seed=seed + 1)
    fd = design.build_fixed_design([[c] for c in cg])
    return fd, np.arange(n, dtype=np.int32), y, Q


def test_auto_below_5000_is_the_dense_exact_run(cb, ctx):
    fd, aid, y, Q = gblup(500, 1500)
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    R1 = mme.AIREML(M, mme.reml_options(algorithm=_lib.CB_VCE_AUTO))
    R1.solve()
    R2 = mme.AIREML(M, mme.reml_options(algorithm=EXACT))
    R2.solve()
    assert R1.history == R2.history and (R1.var_u, R1.var_e) == (R2.var_u, R2.var_e)


@pytest.mark.timeout(900)
def test_auto_direct_regime_matches_oracle(cb, ctx):
    """dim = 5,110 with a dense G^-1: EM first, then AI steps accepted on a
    logL increase, adaptive traces (33..100 probes) — the device run against
    the oracle's run with the same probes and the same rules."""
    fd, aid, y, Q = gblup(5100, 1200)
    M = mme.MixedModel(fd, aid, y, Q.shape[0])
    M.set_qinv_dense(Q)
    R = mme.AIREML(M, mme.reml_options(algorithm=_lib.CB_VCE_AUTO,
                                       solve=mme.solve_options(precision=_lib.CB_PREC_FP64, tol=1e-12)))
    b = R.branch()
    assert (b["first"], b["later"], b["trace"], b["logl_checked"]) == (EM, AI, 1, True)
    Zp = R.probes()
    assert Zp.shape == (M.q, 100)
    R.solve()
    o = orl.reml_run(design.dense_x(fd), aid, y, Q, algorithm="em", later="ai", Zp=Zp,
                     adaptive=(b["min_probes"], b["max_probes"]), logl_check=True, max_iter=20, tol=1e-4)
    print(f"[auto] device {R.iterations} it vu={R.var_u:.8f} ve={R.var_e:.8f}; "
          f"oracle {len(o['hist'])} it vu={o['vu']:.8f} ve={o['ve']:.8f}", flush=True)
    assert R.iterations == len(o["hist"]) and R.converged == o["converged"]
    # Near convergence the AI steps change logL by less than its rounding
    # (log|C| of a 5110 x 5110 factor, ~1e-9 here), so whether a step or
    # its half passes the logL test is decided by rounding, differently for
    # the device Cholesky and numpy's LU: the paths agree to 1e-6 until
    # then and the estimates to the 1e-4 variance-component bar.
    hist, ohist = np.array(R.history), np.array(o["hist"])
    np.testing.assert_allclose(hist[:3], ohist[:3], rtol=1e-6)
    np.testing.assert_allclose(hist, ohist, rtol=1e-4)
    u, ur = R.solution[M.p:], o["sol"][M.p:]
    assert np.abs(u - ur).max() / np.abs(ur).max() < 1e-4


@pytest.fixture(scope="module")
def c4_model(cb, ctx):
    """n = 20,000 GBLUP on m = 50,000 SNPs (MAF U(0.05, 0.5)), h2 = 0.4, 20
    contemporary groups; G and its eigen-bent inverse built on the device."""
    import torch

    from cosmicblup_amd import linalg

    n, m, h2 = 20000, 50000, 0.4
    rows, pitch = synth.bed_on_device(n, m, maf=(0.05, 0.5), missing_rate=0.0, seed=20264203)
    reader = cb.PlinkReader.from_device_rows(ctx, rows.data_ptr(), pitch, n, m)
    gb = cb.GrmBuilder(cb.Options(), ctx=ctx)
    gb.compute(reader, outputs=cb.CB_GRM_OUT_F64)
    G = torch.empty((n, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    gb.copyGrmToDevice(G.data_ptr(), n)
    del gb, reader
    Q = torch.empty_like(G)
    torch.cuda.synchronize()
    linalg.invert_psd_device(G.data_ptr(), n, n, Q.data_ptr(), n, ctx=ctx)
    rng = np.random.default_rng(7)
    # genetic values from 500 QTL dosages decoded on the device
    sel = rows[torch.from_numpy(rng.choice(m, 500, replace=False)).cuda()]
    codes = torch.stack([(sel >> (2 * t)) & 3 for t in range(4)], dim=2).reshape(500, -1)[:, :n]
    Xq = torch.tensor([2.0, 0.0, 1.0, 0.0], dtype=torch.float64, device="cuda")[codes.long()].T
    g = ((Xq - Xq.mean(0)) @ torch.from_numpy(rng.standard_normal(500)).cuda()).cpu().numpy()
    g *= np.sqrt(h2) / g.std()
    grp = rng.integers(0, 20, n)
    y = 10.0 + rng.standard_normal(20)[grp] + g + rng.standard_normal(n) * np.sqrt(1 - h2)
    fd = design.build_fixed_design([[f"cg{int(c):02d}"] for c in grp])
    M = mme.MixedModel(fd, np.arange(n, dtype=np.int32), y, n, ctx=ctx)
    torch.cuda.synchronize()
    M.set_qinv_dense_device(Q.data_ptr(), n)
    del G, Q, rows, sel, codes, Xq
    torch.cuda.empty_cache()
    return M


@pytest.mark.timeout(900)
def test_em_and_ai_share_the_fixed_point_at_n20000(c4_model):
    """The reference runs EM at this size (dense G^-1, dim >= 15000); the AI
    run on the same 64 probes must reach the same variance components."""
    M = c4_model
    so = mme.solve_options(precision=_lib.CB_PREC_MIXED, tol=1e-10)
    Ra = mme.AIREML(M, mme.reml_options(algorithm=AI, n_probes=64, seed=11, tol=1e-7, max_iter=40, solve=so))
    Ra.solve()
    Re = mme.AIREML(M, mme.reml_options(algorithm=EM, n_probes=64, seed=11, tol=1e-8, max_iter=600, solve=so))
    Re.solve()
    print(f"[c4] AI {Ra.iterations} it ({Ra.var_u:.6f}, {Ra.var_e:.6f}); EM {Re.iterations} it "
          f"({Re.var_u:.6f}, {Re.var_e:.6f})", flush=True)
    assert Ra.converged and Re.converged
    assert abs(Re.var_u - Ra.var_u) / Ra.var_u < 1e-4
    assert abs(Re.var_e - Ra.var_e) / Ra.var_e < 1e-4
    # the reference's branch at this size: EM with adaptive probes
    b = mme.AIREML(M, mme.reml_options(algorithm=_lib.CB_VCE_AUTO)).branch()
    assert (b["first"], b["later"], b["trace"]) == (EM, EM, 1)


@pytest.mark.timeout(900)
def test_score_at_n20000_exact_trace_vs_central_fd(c4_model):
    """The data-driven score with the exact trace tr(Q C^uu) (q unit probes,
    in blocks of 2048 right-hand sides) against central finite differences
    (step 1e-4 relative) of the device logL (dense Cholesky of C); and the
    64-probe Hutchinson trace within 10% of the exact one."""
    M = c4_model
    vu, ve = 0.5, 0.65
    so = mme.solve_options(precision=_lib.CB_PREC_MIXED, tol=1e-10)
    Rx = mme.AIREML(M, mme.reml_options(algorithm=AI, trace_exact=True, block_cols=2048, solve=so))
    g_exact, _, tr_exact, _ = Rx.score(vu, ve, want_ai=False)
    hu, he = 1e-4 * vu, 1e-4 * ve
    fd = np.array([(Rx.logL(vu + hu, ve) - Rx.logL(vu - hu, ve)) / (2 * hu),
                   (Rx.logL(vu, ve + he) - Rx.logL(vu, ve - he)) / (2 * he)])
    print(f"[c4] exact-trace score {g_exact}, central FD {fd}, tr {tr_exact:.4f}", flush=True)
    np.testing.assert_allclose(g_exact, fd, rtol=2e-5, atol=1e-5 * np.abs(fd).max())
    R64 = mme.AIREML(M, mme.reml_options(algorithm=AI, n_probes=64, seed=11, solve=so))
    _, _, tr64, _ = R64.score(vu, ve, want_ai=False)
    print(f"[c4] 64-probe trace {tr64:.4f} vs exact {tr_exact:.4f}", flush=True)
    assert abs(tr64 - tr_exact) / tr_exact < 0.1
