"""Benchmark of the GRM hot path on one MI355X.

Workload (the GRM-at-scale configuration): n = 50,000 samples, m = 50,000
SNPs, MAF ~ U(0.01, 0.5) under HWE, 1% MCAR missing calls (mean-imputed
through the missing-call correction), PLINK BED resident in HBM. One step
is one full GrmBuilder::compute: SNP statistics, packing, integer Gram on
INT8 MFMA, FP64 corrections, FP64 + FP32 G written.

Metric: GRM TOPS = m * n * (n + 1) algorithmic integer ops (unique pairs,
MAC = 2) / step time. The roofline object is for the dominant kernel of the
step (timed with HIP events on its launch stream); cpu_baseline is the
reference's GRM arithmetic (centred Z, FP64 Z Z^T through numpy's BLAS) on a
bounded sample, timed on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_INT8_TOPS = 5000.0  # MI355X dense INT8 MFMA, TOP/s
PEAK_HBM_GBS = 8000.0    # MI355X HBM3E, GB/s
PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 MFMA, TFLOP/s
# HBM bytes per k_dense_mfma<float> launch at q = 20,000, r = 65 (the REML
# block product), from the same calibrated FETCH_SIZE / WRITE_SIZE runs:
# 1.75 GB read + 74 MB written
K_DENSE_F32_TRAFFIC_BYTES = 1.824e9
# HBM bytes per launch at n = m = 50,000 with 1% missing, from two rocprofv3
# --pmc runs of this script's GRM leg (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the
# factors a 1 GiB device copy gives):
#   k_gram        86.7 GB read + 30.0 GB written
#   k_miss_cross  76.7 GB read + 20.3 GB written
K_GRAM_TRAFFIC_BYTES = 1.167e11
K_MISS_TRAFFIC_BYTES = 9.70e10
# headline metric: the GRM part is `value` (TOPS on the n = m = 50,000 GRM);
# the REML part (GPU mixed vs GPU FP64 plain iterations/s) is under "reml"
METRIC = "REML iters/sec + GRM TOPS, GPU mixed vs GPU FP64 at \u22641e-6 rel EBV error"


def blas_threads() -> int:
    try:
        import threadpoolctl

        return max(i["num_threads"] for i in threadpoolctl.threadpool_info()) or 1
    except Exception:
        return os.cpu_count() or 1


def cpu_baseline(n_cpu: int = 3072, m_cpu: int = 50000, seed: int = 4) -> dict:
    """The reference's GRM arithmetic on the host cores: the centred,
    mean-imputed Z (missing -> 0, float-rounded 2p) of an n_cpu-sample
    block over all m_cpu SNPs and G_block = Z Z^T as one FP64 GEMM — numpy's
    BLAS standing in for Eigen's blocked GEMM of grm_builder.cpp:351. Z's
    construction is timed with it."""
    import numpy as np

    from cosmicblup_amd import synth
    from oracle import grm as og

    x, _ = synth.simulate_dosages(n_cpu, m_cpu, maf=(0.01, 0.5), missing_rate=0.01, seed=seed)
    bed = synth.encode_bed(x)
    st = og.snp_stats(bed, n_cpu, m_cpu)
    v = st["valid"].astype(bool)
    mu = st["mean_f"].astype(np.float64)[v]
    t0 = time.perf_counter()
    xv = x[:, v]
    Z = np.where(xv >= 0, xv - mu, 0.0)
    G = Z @ Z.T
    dt = time.perf_counter() - t0
    del Z, G, x
    ops = m_cpu * n_cpu * (n_cpu + 1)  # the GPU's op count convention
    return {"value": ops / dt / 1e12, "unit": "TOPS", "cores": blas_threads(), "kind": "port",
            "sample": f"FP64 Z Z^T (numpy BLAS GEMM, stand-in for Eigen's GEMM) of a {n_cpu}-sample block over "
                      f"{m_cpu} SNPs, Z built in the timed region ({dt:.2f} s); the full n = 50,000 GRM is "
                      f"(50000/{n_cpu})^2 such blocks"}


def cpu_reml_c1(n: int = 2000, m: int = 10000, seed: int = 20261203) -> dict:
    """The oracle's dense-exact AI-REML (run_dense_exact_step,
    vce.cpp:2859-3195, the reference's branch below dim 5000) at n = 2000,
    m = 10,000, 20 contemporary groups, on the host cores (numpy LAPACK)."""
    import numpy as np

    from cosmicblup_amd import design, synth
    from oracle import grm as og
    from oracle import reml as orl

    x, _ = synth.simulate_dosages(n, m, maf=(0.05, 0.5), seed=seed)
    G, _, _ = og.grm_vanraden(synth.encode_bed(x), n, m)
    Q = og.invert_psd_matrix(G)
    y, cg = synth.simulate_phenotype(x, n_qtl=500, h2=0.3, n_groups=20, seed=seed + 1)
    X = design.dense_x(design.build_fixed_design([[c] for c in cg]))
    aid = np.arange(n, dtype=np.int32)
    t0 = time.perf_counter()
    o = orl.dense_exact_run(X, aid, y, Q, max_iter=20, tol=1e-4)
    dt = time.perf_counter() - t0
    its = len(o["hist"])
    return {"value": its / dt, "unit": "REML iterations/s", "cores": blas_threads(), "kind": "port",
            "sample": f"oracle dense-exact AI-REML, n = {n}, m = {m}: {its} iterations in {dt:.2f} s"}


def decode_rows_dosage(rows, pitch: int, n: int, snps):
    """Dosage matrix (n x len(snps), float64, missing -> 0) of selected SNP rows
    of a device BED (torch ops; setup only)."""
    import torch

    sel = rows[snps.to(rows.device)]  # k x pitch bytes
    codes = torch.stack([(sel >> (2 * t)) & 3 for t in range(4)], dim=2).reshape(sel.shape[0], -1)[:, :n]
    table = torch.tensor([2.0, 0.0, 1.0, 0.0], dtype=torch.float64, device=rows.device)
    return table[codes.long()].T.contiguous()


def reml_leg(ctx, n: int = 20000, m: int = 50000, k: int = 64, h2: float = 0.4, ngroups: int = 20,
             max_iter: int = 20, plain_iters: int = 0, seed: int = 20264203, cpu: bool = True) -> dict:
    """REML iterations/s, GPU mixed (FP32 block PCG of 64 probes + y with FP64
    refinement) against GPU FP64 plain (FP64 PCG, one right-hand side at a
    time), same AI-REML algorithm and probes."""
    import time

    import numpy as np
    import torch

    import cosmicblup_amd as cb
    from cosmicblup_amd import _lib, design, linalg, mme, synth

    print(f"[bench] REML leg: n={n} m={m} probes={k}", flush=True)
    rows, pitch = synth.bed_on_device(n, m, maf=(0.05, 0.5), missing_rate=0.0, seed=seed)
    reader = cb.PlinkReader.from_device_rows(ctx, rows.data_ptr(), pitch, n, m)
    gb = cb.GrmBuilder(cb.Options(), ctx=ctx)
    gb.compute(reader, outputs=cb.CB_GRM_OUT_F64)
    # G on the device (torch buffer), then G^-1 by the library's eigen-bent
    # inverse (invert_psd_matrix, as the reference's GBLUP path builds it) — set-up
    G = torch.empty((n, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    gb.copyGrmToDevice(G.data_ptr(), n)
    del gb, reader
    Q = torch.empty_like(G)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    linalg.invert_psd_device(G.data_ptr(), n, n, Q.data_ptr(), n, ctx=ctx)
    t_inv = time.perf_counter() - t0
    print(f"[bench] G^-1 (eigen-bent) in {t_inv:.1f} s", flush=True)
    del G
    torch.cuda.synchronize()
    # phenotype: 500 QTL, h2, 20 contemporary groups
    rng = np.random.default_rng(seed)
    qtl = torch.from_numpy(rng.choice(m, 500, replace=False))
    Xq = decode_rows_dosage(rows, pitch, n, qtl)
    del rows
    g = (Xq - Xq.mean(0)) @ torch.from_numpy(rng.standard_normal(500)).cuda()
    g = (g * np.sqrt(h2) / g.std()).cpu().numpy()
    grp = rng.integers(0, ngroups, n)
    y = 10.0 + rng.standard_normal(ngroups)[grp] + g + rng.standard_normal(n) * np.sqrt(1 - h2)
    fd = design.build_fixed_design([[f"cg{int(c):02d}"] for c in grp])
    model = mme.MixedModel(fd, np.arange(n, dtype=np.int32), y, n, ctx=ctx)
    model.set_qinv_dense_device(Q.data_ptr(), n)
    Qh = Q.cpu().numpy() if cpu else None
    del Q
    torch.cuda.empty_cache()

    def run(precision, batch, iters):
        opts = mme.reml_options(n_probes=k, seed=7, max_iter=iters,
                                solve=mme.solve_options(precision=precision, batch_rhs=batch, tol=1e-10))
        R = mme.AIREML(model, opts)
        ctx.synchronize()
        t0 = time.perf_counter()
        R.solve()
        ctx.synchronize()
        return R, time.perf_counter() - t0

    ctx.reset_profile()
    ctx.set_profiling(True)
    Rm, tm = run(_lib.CB_PREC_MIXED, True, max_iter)
    ctx.set_profiling(False)
    print(f"[bench] mixed: {Rm.iterations} iterations in {tm:.2f} s, vu={Rm.var_u:.5f} ve={Rm.var_e:.5f}", flush=True)
    symm_ms, symm_n = ctx.kernel_time("k_dense_mfma_f32")
    # the same run with the dense product of the FP32 recursion on FP16
    # matrix cores (hi/lo halves)
    ctx.reset_profile()
    ctx.set_profiling(True)
    Rh, th = run(_lib.CB_PREC_MIXED_F16, True, max_iter)
    ctx.set_profiling(False)
    print(f"[bench] mixed f16: {Rh.iterations} iterations in {th:.2f} s, vu={Rh.var_u:.5f} ve={Rh.var_e:.5f}",
          flush=True)
    h_ms, h_n = ctx.kernel_time("k_dense_h")
    # the reference's own branch at this size (CB_VCE_AUTO: EM with adaptive
    # Hutchinson probes, dense G^-1 with dim >= 15000), mixed precision
    ra = mme.reml_options(algorithm=_lib.CB_VCE_AUTO, max_iter=max_iter,
                          solve=mme.solve_options(precision=_lib.CB_PREC_MIXED, tol=1e-10))
    Ra = mme.AIREML(model, ra)
    branch = Ra.branch()
    ctx.synchronize()
    t0 = time.perf_counter()
    Ra.solve()
    ctx.synchronize()
    ta = time.perf_counter() - t0
    print(f"[bench] auto (reference branch {branch}): {Ra.iterations} iterations in {ta:.2f} s, "
          f"vu={Ra.var_u:.5f} ve={Ra.var_e:.5f}", flush=True)
    # FP64 plain over the same run as mixed (plain_iters > 0 caps it)
    Rp, tp = run(_lib.CB_PREC_FP64, False, plain_iters or max_iter)
    print(f"[bench] fp64 plain: {Rp.iterations} iterations in {tp:.2f} s", flush=True)
    # FP64 with the same batching: the variance-component reference for the
    # mixed path (same probes, same AI-REML) and a second speed point
    Rb, tb = run(_lib.CB_PREC_FP64, True, max_iter)
    print(f"[bench] fp64 batched: {Rb.iterations} iterations in {tb:.2f} s", flush=True)
    vc_err = max(abs(Rm.var_u - Rb.var_u) / Rb.var_u, abs(Rm.var_e - Rb.var_e) / Rb.var_e)
    vc_err_h = max(abs(Rh.var_u - Rb.var_u) / Rb.var_u, abs(Rh.var_e - Rb.var_e) / Rb.var_e)
    # accuracy of the mixed path against FP64 at identical variances
    opts64 = mme.reml_options(n_probes=k, seed=7, solve=mme.solve_options(precision=_lib.CB_PREC_FP64, tol=1e-12))
    opts32 = mme.reml_options(n_probes=k, seed=7, solve=mme.solve_options(precision=_lib.CB_PREC_MIXED, tol=1e-10))
    g64, _, tr64, s64 = mme.AIREML(model, opts64).score(Rm.var_u, Rm.var_e, want_ai=False)
    g32, _, tr32, s32 = mme.AIREML(model, opts32).score(Rm.var_u, Rm.var_e, want_ai=False)
    opts16 = mme.reml_options(n_probes=k, seed=7, solve=mme.solve_options(precision=_lib.CB_PREC_MIXED_F16, tol=1e-10))
    g16, _, tr16, s16 = mme.AIREML(model, opts16).score(Rm.var_u, Rm.var_e, want_ai=False)
    p = model.p
    ebv_err = float(np.abs(s32[p:] - s64[p:]).max() / np.abs(s64[p:]).max())
    ebv_err_h = float(np.abs(s16[p:] - s64[p:]).max() / np.abs(s64[p:]).max())
    rate_m = Rm.iterations / tm
    rate_p = Rp.iterations / tp
    cpu_leg = None
    if cpu:
        # one FP64 REML iteration of the oracle on the host cores at this
        # size: the EM step with the same 64 probes + y, C factorised densely
        # (the reference's EM branch instead solves each probe by PCG)
        from oracle import reml as orl

        # bounded sample: the first n_s animals' model (their records and
        # block of Q), scaled by the dense factorisation's cube law
        n_s = min(n, 10000)
        Zp = orl.probes(n_s, k, seed=7)
        Xd = design.dense_x(fd)[:n_s]
        Xd = Xd[:, np.abs(Xd).sum(0) > 0]
        Qs = np.ascontiguousarray(Qh[:n_s, :n_s])
        del Qh
        t0 = time.perf_counter()
        orl.reml_eval(Xd, np.arange(n_s, dtype=np.int32), y[:n_s], Qs, Rm.var_u, Rm.var_e, Zp, want_ai=False)
        dt_s = time.perf_counter() - t0
        del Qs
        scale = ((n + fd.p) / (n_s + Xd.shape[1])) ** 3
        cpu_leg = {"value": 1.0 / (dt_s * scale), "unit": "REML iterations/s", "cores": blas_threads(),
                   "kind": "port",
                   "sample": f"one oracle EM iteration (dense LU of C, 64 probes + y) timed on the first {n_s} "
                             f"animals ({dt_s:.1f} s) and scaled by (dim / dim_sample)^3 = {scale:.2f} to "
                             f"n = {n}"}
        print(f"[bench] CPU oracle EM iteration at n={n_s}: {dt_s:.1f} s (x{scale:.2f})", flush=True)
    # the dominant kernel of the mixed run: the FP32 block product Q X_u,
    # 2 q^2 r flops per launch (r = the probes + y)
    dense_ms = symm_ms / max(symm_n, 1)
    dense_tf = 2.0 * n * n * (k + 1) / (dense_ms * 1e-3) / 1e12 if symm_n else None
    roof = {"kernel": "k_dense_mfma_f32", "bound": "mfma", "achieved": dense_tf, "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": dense_tf / PEAK_FP32_TFLOPS if dense_tf else None,
            "traffic": K_DENSE_F32_TRAFFIC_BYTES if (n, k) == (20000, 64) else None}
    return {"n": n, "m": m, "probes": k, "iters_per_s_mixed": rate_m, "iters_per_s_fp64_plain": rate_p,
            "speedup": rate_m / rate_p, "iters_per_s_fp64_batched": Rb.iterations / tb,
            "iterations_mixed": Rm.iterations, "converged": Rm.converged,
            "var_u": Rm.var_u, "var_e": Rm.var_e, "vc_rel_err_mixed_vs_fp64": vc_err,
            "ebv_rel_err_mixed_vs_fp64": ebv_err,
            "trace_rel_err_mixed_vs_fp64": abs(tr32 - tr64) / abs(tr64),
            "k_dense_mfma_f32_avg_ms": symm_ms / max(symm_n, 1),
            "mixed_f16": {"iters_per_s": Rh.iterations / th, "speedup_vs_fp64_plain": Rh.iterations / th / rate_p,
                          "iterations": Rh.iterations, "converged": Rh.converged,
                          "vc_rel_err_vs_fp64": vc_err_h, "ebv_rel_err_vs_fp64": ebv_err_h,
                          "trace_rel_err_vs_fp64": abs(tr16 - tr64) / abs(tr64),
                          "k_dense_h_avg_ms": h_ms / max(h_n, 1)},
            "auto_reference_branch": {"branch": branch, "iterations": Ra.iterations, "converged": Ra.converged,
                                      "seconds": ta, "iters_per_s": Ra.iterations / ta, "var_u": Ra.var_u,
                                      "var_e": Ra.var_e},
            "qinv": "eigen-bent G^-1 (invert_psd_matrix semantics)", "qinv_setup_s": t_inv,
            "roofline": roof, "cpu_baseline": cpu_leg}


def c1_leg(ctx, n: int = 2000, m: int = 10000, h2: float = 0.3, ngroups: int = 20, seed: int = 20261203) -> dict:
    """The reference's CPU-runnable GBLUP case: n=2000, m=10000, no missing,
    G^-1 eigen-bent, AI-REML by the dense-exact step (the branch the reference
    takes below p+q = 5000), from (1, 1) to sum |d sigma^2| < 1e-4."""
    import time

    import numpy as np
    import torch

    import cosmicblup_amd as cb
    from cosmicblup_amd import _lib, design, linalg, mme, synth

    print(f"[bench] C1 leg: n={n} m={m} dense-exact AI-REML", flush=True)
    rows, pitch = synth.bed_on_device(n, m, maf=(0.05, 0.5), missing_rate=0.0, seed=seed)
    reader = cb.PlinkReader.from_device_rows(ctx, rows.data_ptr(), pitch, n, m)
    gb = cb.GrmBuilder(cb.Options(), ctx=ctx)
    gb.compute(reader, outputs=cb.CB_GRM_OUT_F64)
    G = torch.empty((n, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    gb.copyGrmToDevice(G.data_ptr(), n)
    Q = torch.empty_like(G)
    torch.cuda.synchronize()
    linalg.invert_psd_device(G.data_ptr(), n, n, Q.data_ptr(), n, ctx=ctx)
    rng = np.random.default_rng(seed)
    qtl = torch.from_numpy(rng.choice(m, 500, replace=False))
    Xq = decode_rows_dosage(rows, pitch, n, qtl)
    g = (Xq - Xq.mean(0)) @ torch.from_numpy(rng.standard_normal(500)).cuda()
    g = (g * np.sqrt(h2) / g.std()).cpu().numpy()
    grp = rng.integers(0, ngroups, n)
    y = 10.0 + rng.standard_normal(ngroups)[grp] + g + rng.standard_normal(n) * np.sqrt(1 - h2)
    fd = design.build_fixed_design([[f"cg{int(c):02d}"] for c in grp])
    model = mme.MixedModel(fd, np.arange(n, dtype=np.int32), y, n, ctx=ctx)
    model.set_qinv_dense_device(Q.data_ptr(), n)
    del G, Q, rows, gb, reader
    R = mme.AIREML(model, mme.reml_options(algorithm=_lib.CB_VCE_AI_EXACT, max_iter=20, tol=1e-4))
    R.solve()  # warm-up (allocations, library initialisation)
    ctx.synchronize()
    t0 = time.perf_counter()
    R.solve()
    ctx.synchronize()
    dt = time.perf_counter() - t0
    print(f"[bench] C1 dense-exact: {R.iterations} iterations in {dt:.3f} s, vu={R.var_u:.5f} ve={R.var_e:.5f}",
          flush=True)
    return {"n": n, "m": m, "algorithm": "dense-exact AI (vce.cpp:2859-3195)", "iterations": R.iterations,
            "converged": R.converged, "seconds": dt, "iters_per_s": R.iterations / dt, "var_u": R.var_u,
            "var_e": R.var_e, "se_var_u": R.getSEVarsU()[0], "se_var_e": R.getSEVarE()}


def c5_leg(ctx, n_inst: int = 96, m: int = 20000, n_lo: int = 500, n_hi: int = 8000, k: int = 32,
           seed: int = 20265203, workers: int = 4) -> dict:
    """Batched uneven instances: n log-uniform in [n_lo, n_hi], m SNPs,
    MAF U(0.05, 0.5); even instances single-trait (AI-REML, mixed batched
    PCG with k probes + y), odd instances 3-trait (r_g ~ U(-0.5, 0.8), h2 ~
    U(0.1, 0.6), r_e = 0.2; projected eigen-rotation EM, then the Kronecker
    MME solution by the canonical transformation and PCG). Per instance:
    GRM and G^-1 on the device (set-up, untimed), then the timed estimation,
    serially. Reports REML (or EM) iterations/s per size bucket. Then every
    instance is served again together: each of `workers` host threads
    drives its own context (stream) through its share (balanced on n^2,
    largest first), and the batch's wall time is compared with the serial
    sum; the estimates must be bit-identical to the serial ones."""
    import math
    import time

    import numpy as np
    import torch

    import cosmicblup_amd as cb
    from cosmicblup_amd import _lib, design, linalg, mme, multitrait, synth

    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(n_lo), math.log(n_hi), n_inst)).astype(int)
    print(f"[bench] C5 leg: {n_inst} instances, n in [{sizes.min()}, {sizes.max()}], m={m}", flush=True)
    # instances to serving contexts: largest n^2 first, each to the least
    # loaded worker
    ctxs = [ctx] + [_lib.Context(0) for _ in range(workers - 1)]
    load = [0.0] * workers
    owner = {}
    for idx in sorted(range(n_inst), key=lambda i: -int(sizes[i])):
        w = min(range(workers), key=lambda j: load[j])
        owner[idx] = w
        load[w] += float(sizes[idx]) ** 2
    served = []  # (worker, idx, job, serial result); job() reruns the timed estimation
    recs = []
    keep = []  # device inputs the served jobs read again
    for idx, n in enumerate(sizes):
        n = int(n)
        rows, pitch = synth.bed_on_device(n, m, maf=(0.05, 0.5), missing_rate=0.0, seed=seed + 1 + idx)
        reader = cb.PlinkReader.from_device_rows(ctx, rows.data_ptr(), pitch, n, m)
        gb = cb.GrmBuilder(cb.Options(), ctx=ctx)
        gb.compute(reader, outputs=cb.CB_GRM_OUT_F64)
        G = torch.empty((n, n), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        gb.copyGrmToDevice(G.data_ptr(), n)
        Q = torch.empty_like(G)
        torch.cuda.synchronize()
        linalg.invert_psd_device(G.data_ptr(), n, n, Q.data_ptr(), n, ctx=ctx)
        grp = rng.integers(0, 10, n)
        fd = design.build_fixed_design([[f"cg{int(c)}"] for c in grp])
        qtl = torch.from_numpy(rng.choice(m, 200, replace=False))
        Xq = decode_rows_dosage(rows, pitch, n, qtl)
        Xc = (Xq - Xq.mean(0)).cpu().numpy()
        del rows, gb, reader
        mctx = ctxs[owner[idx]]
        if idx % 2 == 0:
            h2 = float(rng.uniform(0.1, 0.6))
            g = Xc @ rng.standard_normal(200)
            g *= math.sqrt(h2) / g.std()
            y = 10.0 + rng.standard_normal(10)[grp] + g + rng.standard_normal(n) * math.sqrt(1 - h2)
            ctx.synchronize()
            model = mme.MixedModel(fd, np.arange(n, dtype=np.int32), y, n, ctx=mctx)
            model.set_qinv_dense_device(Q.data_ptr(), n)
            ropts = mme.reml_options(n_probes=k, seed=idx, max_iter=20, tol=1e-4, solve=mme.solve_options(tol=1e-8))

            def job(model=model, ropts=ropts):
                R = mme.AIREML(model, ropts)
                R.solve()
                model.ctx.synchronize()
                return R, (R.iterations, R.var_u, R.var_e, R.solution.tobytes())
            del G, Q
        else:
            t = 3
            h2 = rng.uniform(0.1, 0.6, t)
            rg = rng.uniform(-0.5, 0.8, 3)
            Cg = np.array([[1, rg[0], rg[1]], [rg[0], 1, rg[2]], [rg[1], rg[2], 1]])
            w, V = np.linalg.eigh(Cg)
            Cg = (V * np.maximum(w, 0.05)) @ V.T  # nearest valid correlation-like matrix
            d = np.sqrt(np.diag(Cg))
            Cg = Cg / np.outer(d, d)
            Vg = Cg * np.sqrt(np.outer(h2, h2))
            Ve = (0.2 + 0.8 * np.eye(t)) * np.sqrt(np.outer(1 - h2, 1 - h2))
            B = Xc @ rng.standard_normal((200, t))
            B = (B - B.mean(0)) / B.std(0)
            U = B @ np.linalg.cholesky(Vg + 1e-9 * np.eye(t)).T
            E = rng.standard_normal((n, t)) @ np.linalg.cholesky(Ve).T
            X = design.dense_x(fd)
            Y = X @ rng.standard_normal((X.shape[1], t)) + U + E
            keep.append((G, Q))

            def job(Y=Y, X=X, G=G, Q=Q, mctx=mctx):
                v = multitrait.MvEigenVce(multitrait.MvEigenVce.Options(max_iter=500, tol=1e-4), ctx=mctx)
                v.prepare(Y, X, G)
                v.runNullModel()
                beta, u = multitrait.solveMultiTrait(Y, X, Q, v.getVg(), v.getVe(), ctx=mctx)
                mctx.synchronize()
                return v, (v.iterations, v.getVg().tobytes(), v.getVe().tobytes(), beta.tobytes(), u.tobytes())
        ctx.synchronize()
        mctx.synchronize()
        t0 = time.perf_counter()
        est, bits = job()
        dt = time.perf_counter() - t0
        traits = 1 if idx % 2 == 0 else 3
        recs.append({"n": n, "traits": traits, "iterations": est.iterations, "seconds": dt,
                     "converged": bool(est.converged)})
        served.append((owner[idx], idx, job, bits))
        torch.cuda.empty_cache()
        print(f"[bench] C5 instance {idx + 1}/{n_inst}: n={n} traits={traits} "
              f"{recs[-1]['iterations']} iterations in {dt:.3f} s", flush=True)
    buckets = [(500, 1000), (1000, 2000), (2000, 4000), (4000, 8001)]
    out = {"instances": n_inst, "m": m, "probes": k, "per_bucket": []}
    for lo, hi in buckets:
        for tr in (1, 3):
            sel = [r for r in recs if lo <= r["n"] < hi and r["traits"] == tr]
            if not sel:
                continue
            it = sum(r["iterations"] for r in sel)
            sec = sum(r["seconds"] for r in sel)
            out["per_bucket"].append({"n_range": [lo, hi], "traits": tr, "instances": len(sel),
                                      "iterations": it, "seconds": sec, "iters_per_s": it / sec,
                                      "converged": sum(r["converged"] for r in sel)})
    out["total_seconds"] = sum(r["seconds"] for r in recs)

    def serve(w):
        return [(idx, job()[1]) for wk, idx, job, _ in served if wk == w]

    for c in ctxs:
        c.synchronize()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(workers) as ex:
        got = dict(sum(ex.map(serve, range(workers)), []))
    wall = time.perf_counter() - t0
    serial = out["total_seconds"]
    its = sum(r["iterations"] for r in recs)
    out["served"] = {
        "instances": len(served), "workers": workers, "iterations": its, "wall_seconds": wall,
        "serial_seconds": serial, "speedup_vs_serial": serial / wall,
        "identical_to_serial": all(got[idx] == ref for _, idx, _, ref in served),
        "differing_instances": [idx for _, idx, _, ref in served if got[idx] != ref]}
    print(f"[bench] C5 served: {len(served)} instances on {workers} streams in {wall:.3f} s "
          f"(serial {serial:.3f} s), identical {out['served']['identical_to_serial']}", flush=True)
    del served, keep
    return out


def cpu_pcg(fd, rec, y, q, A, lam, iters: int = 20) -> dict:
    """MatrixFreePCG (matrix_free_pcg.h:20-66) iterations on the host: the
    PBLUP MME C = W'W + lambda A^-1 as one scipy CSR (W = [X Z] from the
    records), Jacobi preconditioner, one thread (the reference's default
    threads = 1, config.h:39); `iters` iterations timed."""
    import numpy as np
    import scipy.sparse as sp

    n = len(y)
    X = sp.csr_matrix((fd.val, fd.col, fd.ptr), shape=(n, fd.p))
    Z = sp.csr_matrix((np.ones(n), np.asarray(rec, dtype=np.int64), np.arange(n + 1)), shape=(n, q))
    W = sp.hstack([X, Z]).tocsr()
    Ai = sp.csr_matrix((A.val, A.col, A.ptr), shape=(q, q))
    Cm = (W.T @ W + sp.block_diag([sp.csr_matrix((fd.p, fd.p)), lam * Ai])).tocsr()
    b = W.T @ np.asarray(y, dtype=np.float64)
    dinv = 1.0 / Cm.diagonal()
    x = np.zeros_like(b)
    r = b.copy()
    z = dinv * r
    pv = z.copy()
    rz = r @ z
    t0 = time.perf_counter()
    for _ in range(iters):
        Ap = Cm @ pv
        alpha = rz / (pv @ Ap)
        x += alpha * pv
        r -= alpha * Ap
        z = dinv * r
        rz_new = r @ z
        pv = z + (rz_new / rz) * pv
        rz = rz_new
    dt = time.perf_counter() - t0
    return {"value": iters / dt, "unit": "PCG iterations/s", "cores": 1, "kind": "port",
            "sample": f"{iters} Jacobi-PCG iterations on the {Cm.shape[0]}-dim PBLUP MME (scipy CSR, "
                      f"{Cm.nnz} nonzeros), {dt:.2f} s"}


def pcg_leg(ctx, n_gen: int = 10, per_gen: int = 100000, n_rec: int = 600000, n_groups: int = 2000,
            h2: float = 0.25, n_geno: int = 20000, m_geno: int = 40000, tol: float = 1e-10,
            tol_ebv: float = 1e-12, seed: int = 20263203, cpu: bool = True) -> dict:
    """Sparse-MME PCG at the 1M-animal pedigree shape: A^-1 by Henderson's
    rules with inbreeding (host builder), 600k records, 2000 contemporary
    groups; PBLUP (A^-1) and ssGBLUP (H^-1 with the last n_geno animals
    genotyped at m_geno SNPs, set up on the device). One right-hand side per
    model solved to relative residual `tol` by GPU FP64 PCG and by the mixed
    FP32/FP64-replacement PCG."""
    import time

    import numpy as np
    import torch

    import cosmicblup_amd as cb
    from cosmicblup_amd import _lib, design, mme, ssgblup, synth
    from cosmicblup_amd.pedigree import invertA_henderson

    print(f"[bench] PCG leg: pedigree {n_gen}x{per_gen}, {n_rec} records, {n_groups} groups, "
          f"{n_geno} genotyped x {m_geno} SNPs", flush=True)
    dam, sire = synth.simulate_pedigree(n_gen, per_gen, unknown=0.05, seed=seed)
    n = dam.size
    t0 = time.perf_counter()
    A = invertA_henderson(dam, sire)
    t_ainv = time.perf_counter() - t0
    rec, y, grp = synth.simulate_pedigree_phenotype(dam, sire, A.dii, n_rec, h2=h2, n_groups=n_groups,
                                                    seed=seed + 1)
    width = len(str(n_groups))
    fd = design.build_fixed_design([[f"cg{int(c):0{width}d}"] for c in grp])
    lam = (1.0 - h2) / h2
    out = {"animals": int(n), "records": n_rec, "groups": n_groups, "ainv_nnz": int(A.ptr[-1]),
           "ainv_build_s": t_ainv, "lambda": lam, "tol": tol}

    def run_solves(model, tag):
        b = model.rhs()
        dim = b.size
        dB = torch.from_numpy(b).cuda()
        dX = torch.empty_like(dB)
        torch.cuda.synchronize()
        # reference solution: FP64 to 1e-13
        x_ref, it_ref, _ = model.solve(lam, b, mme.solve_options(precision=_lib.CB_PREC_FP64, tol=1e-13,
                                                                 max_iter=20000))
        p = model.p
        scale = np.abs(x_ref[p:]).max()
        res = {"fp64_1e-13_iterations": it_ref}
        # the reference's default tolerance (timed as the headline) and the
        # one that meets the 1e-6 EBV bar against the 1e-13 solution
        for t in (tol, tol_ebv):
            sols = {}
            key = "" if t == tol else f"_tol_{t:g}"
            for name, prec in (("fp64", _lib.CB_PREC_FP64), ("mixed", _lib.CB_PREC_MIXED)):
                opts = mme.solve_options(precision=prec, tol=t, max_iter=20000)
                model.solve(lam, b, opts)  # warm-up: preconditioner set-up, allocations
                # timed: right-hand side and solution resident on the device
                ctx.synchronize()
                t0 = time.perf_counter()
                it, rel = model.solve_device(lam, dB.data_ptr(), dim, dX.data_ptr(), dim, 1, opts)
                ctx.synchronize()
                dt = time.perf_counter() - t0
                sols[name] = dX.cpu().numpy().copy()
                entry = {"iterations": it, "seconds": dt, "iters_per_s": it / dt, "rel_residual": rel}
                if t == tol:
                    # the same solve from and to host arrays (PCIe and staging included)
                    t0 = time.perf_counter()
                    model.solve(lam, b, opts)
                    entry["seconds_host_arrays"] = time.perf_counter() - t0
                res[name + key] = entry
                print(f"[bench] {tag} PCG {name} tol {t:g}: {it} iterations in {dt:.4f} s on the device "
                      f"(rel {rel:.2e})", flush=True)
            # both stop at relative residual t, so their EBVs differ by up to
            # cond(C) * t; against the 1e-13 solution each shows its own error
            res["ebv_rel_err_mixed_vs_fp64" + key] = float(np.abs(sols["mixed"][p:] - sols["fp64"][p:]).max() / scale)
            res["ebv_rel_err_vs_fp64_1e-13" + key] = {k: float(np.abs(v[p:] - x_ref[p:]).max() / scale)
                                                    for k, v in sols.items()}
        res["time_speedup_mixed_vs_fp64"] = res["fp64"]["seconds"] / res["mixed"]["seconds"]
        res["tol_for_1e-6_ebv"] = tol_ebv
        if tag == "PBLUP":
            # the dominant kernel of an FP64 iteration, timed with HIP events in
            # a solve of its own: the pass over W'W's animal rows that also
            # applies lambda A^-1. Algorithmic bytes per launch: 12 per nonzero
            # (FP64 value + int32 index) of A^-1 and of the animal rows (3 per
            # record), two row pointers, x and y once per row
            opts = mme.solve_options(precision=_lib.CB_PREC_FP64, tol=tol, max_iter=20000)
            ctx.reset_profile()
            ctx.set_profiling(True)
            model.solve_device(lam, dB.data_ptr(), dim, dX.data_ptr(), dim, 1, opts)
            ctx.set_profiling(False)
            ms, cnt = ctx.kernel_time("k_csr_spmm_f64")
            nbytes = 12.0 * (int(A.ptr[-1]) + 3 * n_rec) + 32.0 * n
            gbs = nbytes / (ms / cnt * 1e-3) / 1e9 if cnt else None
            res["roofline"] = {"kernel": "k_csr2_spmm (W'W animal rows + lambda A^-1)", "bound": "hbm",
                               "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                               "frac": gbs / PEAK_HBM_GBS if gbs else None, "traffic": None,
                               "avg_us": ms / cnt * 1e3 if cnt else None, "bytes_per_launch": nbytes}
        return res

    model = mme.MixedModel(fd, rec, y, n, ctx=ctx)
    model.set_qinv_csr(A.ptr, A.col, A.val)
    out["pblup"] = run_solves(model, "PBLUP")
    if cpu:
        out["pblup"]["cpu_baseline"] = cpu_pcg(fd, rec, y, n, A, lam)
    if n_geno > 0:
        # genotyped: the last n_geno animals (youngest generations)
        gmap = np.arange(n - n_geno, n, dtype=np.int32)
        rows, pitch = synth.bed_on_device(n_geno, m_geno, maf=(0.01, 0.5), missing_rate=0.0, seed=seed + 2)
        reader = cb.PlinkReader.from_device_rows(ctx, rows.data_ptr(), pitch, n_geno, m_geno)
        gb = cb.GrmBuilder(cb.Options(), ctx=ctx)
        gb.compute(reader, outputs=cb.CB_GRM_OUT_F64)
        G = torch.empty((n_geno, n_geno), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        gb.copyGrmToDevice(G.data_ptr(), n_geno)
        del gb, reader, rows
        t0 = time.perf_counter()
        D, _, st = ssgblup.build_hinv_blocks(dam, sire, gmap, G, ctx=ctx)
        t_h = time.perf_counter() - t0
        del G
        print(f"[bench] H^-1 blocks in {t_h:.1f} s (tuneG a={st['a']:.4f} b={st['b']:.4f})", flush=True)
        model.set_qinv_split(A.ptr, A.col, A.val, gmap, D.data_ptr(), n_geno)
        del D
        torch.cuda.empty_cache()
        out["ssgblup"] = run_solves(model, "ssGBLUP")
        out["ssgblup"].update({"genotyped": n_geno, "snps": m_geno, "hinv_setup_s": t_h})
    return out


def int8_vendor(n: int, m: int, reps: int = 5) -> dict:
    """The vendor INT8 GEMM (torch._int_mm: rocBLAS / hipBLASLt, int8 in,
    int32 out) on the GRM's shape, X (n x m) times X' (m x n), for a
    reference point beside k_gram. It forms all n^2 entries (2 n^2 m ops);
    k_gram forms the lower triangle (m n (n + 1) ops) from 2-bit codes and
    writes G in FP64 and FP32 in its epilogue."""
    import torch

    X = torch.randint(0, 3, (n, m), dtype=torch.int8, device="cuda")
    for _ in range(2):
        Y = torch._int_mm(X, X.t())  # X' as the column-major view of X
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        Y = torch._int_mm(X, X.t())
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del X, Y
    torch.cuda.empty_cache()
    ops = 2.0 * n * n * m
    return {"n": n, "m": m, "vendor_int8_gemm": {"ms": ms, "tops": ops / ms / 1e9, "ops": ops,
                                                 "frac_int8_peak": ops / ms / 1e9 / PEAK_INT8_TOPS}}


def dense_vs_vendor(ctx, q: int = 20000, r: int = 65, reps: int = 20) -> dict:
    """The REML block product Q X (q x q by q x r) by the library's kernels
    (k_dense_mfma FP32 and FP64, k_dense_h on FP16 halves, each with its
    split-K reduction; synthetic operands, the library's split, HIP events on
    its stream) beside
    torch.matmul of the same shape, which runs the vendor GEMM (rocBLAS /
    hipBLASLt): FP32, FP64, and FP16 (one FP16 copy of Q, not the two halves
    the library's FP16 path reads)."""
    import ctypes as C

    import torch

    from cosmicblup_amd import _lib

    flops = 2.0 * q * q * r
    out = {"q": q, "r": r, "library": {}, "vendor": {}}
    for name, elem in (("fp32", 4), ("fp64", 8), ("fp16_halves", 2)):
        tb, ks = C.c_int(), C.c_int()
        _lib.check(_lib.lib().cb_tune_dense_lookup(elem, q, r, C.byref(tb), C.byref(ks)))
        ms = C.c_double()
        _lib.check(_lib.lib().cb_tune_dense_time(ctx.handle, elem, q, r, tb.value, ks.value, reps, C.byref(ms)))
        out["library"][name] = {"ms": ms.value, "tflops": flops / ms.value / 1e9}
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64), ("fp16", torch.float16)):
        A = torch.randn(q, q, device="cuda", dtype=torch.float32).to(dt)
        X = torch.randn(q, r, device="cuda", dtype=torch.float32).to(dt)
        for _ in range(3):
            Y = A @ X
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            Y = A @ X
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out["vendor"][name] = {"ms": ms, "tflops": flops / ms / 1e9}
        del A, X, Y
        torch.cuda.empty_cache()
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=50000)
    ap.add_argument("--m", type=int, default=50000)
    ap.add_argument("--missing", type=float, default=0.01)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-reml", action="store_true")
    ap.add_argument("--reml-n", type=int, default=20000)
    ap.add_argument("--reml-m", type=int, default=50000)
    ap.add_argument("--plain-iters", type=int, default=0,
                    help="cap the FP64 plain REML run (0: the same max_iter as the mixed run)")
    ap.add_argument("--reml-only", action="store_true",
                    help="only run the REML leg and print its JSON object")
    ap.add_argument("--no-pcg", action="store_true", help="skip the 1M-animal PBLUP PCG leg")
    ap.add_argument("--no-c1", action="store_true", help="skip the n=2000 dense-exact REML leg")
    ap.add_argument("--no-c5", action="store_true", help="skip the 96-instance batched workload")
    ap.add_argument("--c5-only", action="store_true", help="only run the 96-instance workload")
    ap.add_argument("--c5-workers", type=int, default=8, help="contexts (streams and host threads) serving C5")
    ap.add_argument("--graphs", action="store_true",
                    help="replay the PCG iteration blocks as captured HIP graphs (cb_context_set_graphs)")
    ap.add_argument("--pcg-only", action="store_true",
                    help="only run the PBLUP PCG leg and print its JSON object")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a plain device copy of this many bytes (HBM counter calibration)")
    ap.add_argument("--int8-vs-vendor", action="store_true",
                    help="only time the vendor INT8 GEMM (torch._int_mm) on the GRM's n x m by m x n shape")
    ap.add_argument("--dense-vs-vendor", action="store_true",
                    help="only time the REML block product (q = 20,000, 65 columns) by the library's kernels "
                         "and by torch.matmul (the vendor GEMM) in FP32, FP64 and FP16")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: there is no multi-GPU path; one GPU's share "
              "(the GRM tiles, the probe columns of a REML block) is the unit of work", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_bytes:
        # reference point for FETCH_SIZE / WRITE_SIZE: one read and one write
        # of copy_bytes each per copy launch
        a = torch.ones(args.copy_bytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.copy_bytes, "launches": 3}), flush=True)
        return

    import cosmicblup_amd as cb
    from cosmicblup_amd import _lib, synth

    if args.int8_vs_vendor:
        print(json.dumps(int8_vendor(args.n, args.m)), flush=True)
        return
    if args.dense_vs_vendor:
        print(json.dumps(dense_vs_vendor(_lib.Context(0))), flush=True)
        return
    if args.reml_only:
        ctx = _lib.Context(0)
        ctx.set_graphs(args.graphs)
        print(json.dumps(reml_leg(ctx, n=args.reml_n, m=args.reml_m, plain_iters=args.plain_iters)), flush=True)
        return
    if args.pcg_only:
        ctx = _lib.Context(0)
        ctx.set_graphs(args.graphs)
        print(json.dumps(pcg_leg(ctx)), flush=True)
        return
    if args.c5_only:
        ctx = _lib.Context(0)
        print(json.dumps(c5_leg(ctx, workers=args.c5_workers)), flush=True)
        return

    n, m = args.n, args.m
    print(f"[bench] generating BED n={n} m={m} missing={args.missing} on device", flush=True)
    rows, pitch = synth.bed_on_device(n, m, maf=(0.01, 0.5), missing_rate=args.missing, seed=20262203)
    torch.cuda.synchronize()
    ctx = _lib.Context(0)
    ctx.set_graphs(args.graphs)
    reader = cb.PlinkReader.from_device_rows(ctx, rows.data_ptr(), pitch, n, m)
    del rows
    torch.cuda.empty_cache()
    gb = cb.GrmBuilder(cb.Options(), ctx=ctx)
    outputs = cb.CB_GRM_OUT_F64 | cb.CB_GRM_OUT_F32
    for w in range(args.warmup):
        gb.compute(reader, outputs=outputs)
        print(f"[bench] warmup {w + 1}/{args.warmup} done", flush=True)
    ctx.synchronize()
    ctx.reset_profile()
    ctx.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        gb.compute(reader, outputs=outputs)
        print(f"[bench] step {s + 1}/{args.steps} done", flush=True)
    ctx.synchronize()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ctx.set_profiling(False)
    ops = float(m) * n * (n + 1)
    kern = {}
    for k in ("k_snp_stats", "k_pack", "k_row_sums", "k_miss_count", "k_miss_scan", "k_miss_emit", "k_sample_sq",
              "k_encode_iblock", "k_group_entries", "k_miss_cross", "k_gram"):
        ms, cnt = ctx.kernel_time(k)
        if cnt:
            kern[k] = ms / cnt
    c2 = (n, m, args.missing) == (50000, 50000, 0.01)
    gram_ms = kern["k_gram"]
    achieved = ops / (gram_ms * 1e-3) / 1e12
    gram_roof = {"kernel": "k_gram", "bound": "mfma", "achieved": achieved, "peak": PEAK_INT8_TOPS, "unit": "TOPS",
                 "frac": achieved / PEAK_INT8_TOPS, "traffic": K_GRAM_TRAFFIC_BYTES if c2 else None}
    rooflines = [gram_roof]
    if "k_miss_cross" in kern:
        # algorithmic bytes: the packed genotypes read once (n m / 4) and the
        # n x n FP64 correction E written once
        miss_bytes = n * m / 4 + 8.0 * n * n
        gbs = miss_bytes / (kern["k_miss_cross"] * 1e-3) / 1e9
        rooflines.append({"kernel": "k_miss_cross", "bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS,
                          "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
                          "traffic": K_MISS_TRAFFIC_BYTES if c2 else None})
    # the dominant kernel's roofline is the headline one
    rooflines.sort(key=lambda r: -kern[r["kernel"]])
    kern_sum = sum(kern.values())
    out = {
        "metric": METRIC,
        "value": ops / dt / 1e12,
        "unit": "TOPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": f"GRM VanRaden n={n} m={m} MAF U(0.01,0.5) HWE, "
                               f"{args.missing:.0%} MCAR missing, G in FP64+FP32",
                   "n": n, "m": m, "missing": args.missing},
        "roofline": rooflines[0],
        "roofline_other_kernels": rooflines[1:],
        # the whole step against the INT8 MFMA peak
        "step_frac_int8_peak": ops / dt / 1e12 / PEAK_INT8_TOPS,
        "kernel_ms": kern,
        # per step: the kernels' own time and what lies outside them (host
        # statistics, list building, copies, launch gaps)
        "step_breakdown_ms": {"kernels": kern_sum, "outside_kernels": dt * 1e3 - kern_sum},
    }
    del gb, reader
    torch.cuda.empty_cache()
    if not args.no_reml:
        out["reml"] = reml_leg(ctx, n=args.reml_n, m=args.reml_m, plain_iters=args.plain_iters, cpu=not args.no_cpu)
        torch.cuda.empty_cache()
    if not args.no_pcg:
        out["pblup_pcg"] = pcg_leg(ctx, cpu=not args.no_cpu)
        torch.cuda.empty_cache()
    if not args.no_c1:
        out["c1_dense_exact"] = c1_leg(ctx)
        if not args.no_cpu:
            print("[bench] timing the CPU oracle dense-exact REML (n=2000)", flush=True)
            out["c1_dense_exact"]["cpu_baseline"] = cpu_reml_c1()
    if not args.no_c5:
        torch.cuda.empty_cache()
        out["c5_batched"] = c5_leg(ctx, workers=args.c5_workers)
    if not args.no_cpu:
        torch.cuda.empty_cache()
        print("[bench] timing the CPU oracle sample", flush=True)
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
