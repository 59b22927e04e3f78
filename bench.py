"""Benchmark of the GRM hot path on one MI355X.

Workload (the GRM-at-scale configuration): n = 50,000 samples, m = 50,000
SNPs, MAF ~ U(0.01, 0.5) under HWE, 1% MCAR missing calls (mean-imputed
through the missing-call correction), PLINK BED resident in HBM. One step
is one full GrmBuilder::compute: SNP statistics, packing, integer Gram on
INT8 MFMA, FP64 corrections, FP64 + FP32 G written.

Metric: GRM TOPS = m * n * (n + 1) algorithmic integer ops (unique pairs,
MAC = 2) / step time. The roofline object is for the dominant kernel
(k_gram) timed with HIP events on its launch stream; cpu_baseline is the
oracle's OpenMP FP64 GRM on a bounded sub-block, timed on the host cores.
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


def cpu_baseline(n_cpu: int = 1536, m_cpu: int = 50000, seed: int = 4) -> dict:
    """Oracle FP64 GRM (reference algorithm: centred Z, Z Z^T) on an
    n_cpu x n_cpu sub-block over m_cpu SNPs, all host cores."""
    import numpy as np

    from cosmicblup_amd import synth
    from oracle import grm as og

    x, _ = synth.simulate_dosages(n_cpu, m_cpu, maf=(0.01, 0.5), missing_rate=0.01, seed=seed)
    bed = synth.encode_bed(x)
    del x
    st = og.snp_stats(bed, n_cpu, m_cpu)
    t0 = time.perf_counter()
    og.grm_block_raw(bed, n_cpu, m_cpu, st, 0, n_cpu, 0, n_cpu)
    dt = time.perf_counter() - t0
    # same op count convention as the GPU: unique pairs incl. diagonal, MAC = 2
    ops = m_cpu * n_cpu * (n_cpu + 1)
    return {"value": ops / dt / 1e12, "unit": "TOPS", "cores": og.num_threads(), "kind": "port",
            "sample": f"FP64 Z Z^T of a {n_cpu}x{n_cpu} sample block over {m_cpu} SNPs "
                      f"({dt:.1f} s; unique-pair op count m*n(n+1))"}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=50000)
    ap.add_argument("--m", type=int, default=50000)
    ap.add_argument("--missing", type=float, default=0.01)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(one GPU's GRM is the unit; see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)

    import torch

    import cosmicblup_amd as cb
    from cosmicblup_amd import _lib, synth

    n, m = args.n, args.m
    print(f"[bench] generating BED n={n} m={m} missing={args.missing} on device", flush=True)
    rows, pitch = synth.bed_on_device(n, m, maf=(0.01, 0.5), missing_rate=args.missing, seed=20262203)
    torch.cuda.synchronize()
    ctx = _lib.Context(0)
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
    for k in ("k_snp_stats", "k_pack", "k_row_sums", "k_miss_snp_lists", "k_miss_sample_lists",
              "k_miss_f", "k_miss_cross", "k_gram"):
        ms, cnt = ctx.kernel_time(k)
        if cnt:
            kern[k] = ms / cnt
    gram_ms = kern["k_gram"]
    achieved = ops / (gram_ms * 1e-3) / 1e12
    out = {
        "metric": "GRM TOPS (n=50k, m=50k, 1% missing; REML iters/sec reported separately)",
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
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_INT8_TOPS, "unit": "TOPS",
                     "frac": achieved / PEAK_INT8_TOPS, "traffic": None},
        "kernel_ms": kern,
    }
    if not args.no_cpu:
        del gb, reader
        torch.cuda.empty_cache()
        print("[bench] timing the CPU oracle sample", flush=True)
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
