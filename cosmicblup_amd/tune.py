"""Autotuner for the split-K shape of the dense operator kernel.

`k_dense_mfma` computes Q X for a dense q x q block Q (G^-1, or the genotyped
block of H^-1) and an r-column block X. Its grid is (row tiles of 128) x
(k splits); a split's partial products are summed in a fixed order by
`k_dense_reduce`. How many splits pay off depends on q (row tiles fill the
chip at large q only), on r and on the element size, so the library reads
the split from a table, `csrc/cb_dense_tuned.h`, that this tuner writes:

    python -m cosmicblup_amd.tune            # on an MI355X; rewrites the table
    python -m cosmicblup_amd.tune --dry-run  # print the table only

For each element size (FP16 halves of the CB_PREC_MIXED_F16 product, FP32,
FP64), each representative q and r, every (target blocks, minimum k span)
candidate is timed through
`cb_tune_dense_time` (HIP events on the library's stream, synthetic
operands) in three interleaved passes, and the fastest (best of its three)
is kept when it beats the built-in split by more than 5%. A table entry
covers q up to the geometric midpoint to the next representative q and r up
to the next r bucket. `--elems 2` tunes one element size and keeps the
table's rows for the others. The library is rebuilt with the new table
(`__graft_entry__.build()`).

The split changes which partial products are summed together, so results
depend on the table in the last bits; for a given table they are
deterministic (the table is a build input, never consulted at run time
against timings)."""
from __future__ import annotations

import argparse
import ctypes as C
import math
import re
import time
from pathlib import Path

from . import _lib

TABLE = Path(__file__).resolve().parent / "csrc" / "cb_dense_tuned.h"
Q_POINTS = (600, 1200, 2500, 5000, 10000, 20000, 40000)
R_BUCKETS = ((5, 8), (17, 24), (33, 48), (65, 80))  # (r timed, r_max of the entry)
TARGETS = (512, 1024, 2048, 4096, 8192)
KSPANS = (32, 64, 128, 256, 512)
KSPANS_F16 = (192, 384, 768, 1536, 3072)  # k_dense_h runs k in steps of 6 x 32
ELEMS = (2, 4, 8)
DEFAULT = (2048, 256)
ROUNDS = 3       # interleaved timing passes per shape
MARGIN = 0.05    # a candidate must beat the default by this fraction


def time_shape(ctx: _lib.Context, elem: int, q: int, r: int, target: int, kspan: int, reps: int) -> float:
    ms = C.c_double()
    _lib.check(_lib.lib().cb_tune_dense_time(ctx.handle, elem, q, r, target, kspan, reps, C.byref(ms)))
    return ms.value


def lookup(elem: int, q: int, r: int) -> tuple[int, int]:
    tb, mk = C.c_int(), C.c_int()
    _lib.check(_lib.lib().cb_tune_dense_lookup(elem, q, r, C.byref(tb), C.byref(mk)))
    return tb.value, mk.value


def reps_for(elem: int, q: int, r: int) -> int:
    # about 20 ms of work per candidate at the FP32 MFMA rate, 3..50 launches
    flops = 2.0 * q * q * r
    per = flops / (60e12 if elem == 4 else 30e12)
    return int(min(50, max(3, 0.02 / max(per, 1e-6))))


def tune(ctx: _lib.Context, log=print, elems=ELEMS) -> list[dict]:
    entries = []
    for elem in elems:
        for qi, q in enumerate(Q_POINTS):
            q_max = int(math.sqrt(q * Q_POINTS[qi + 1])) if qi + 1 < len(Q_POINTS) else 1 << 40
            for r, r_max in R_BUCKETS:
                reps = reps_for(elem, q, r)
                spans = KSPANS_F16 if elem == 2 else KSPANS
                cands = [DEFAULT] + [(tb, mk) for tb in TARGETS for mk in spans if (tb, mk) != DEFAULT]
                # ROUNDS interleaved passes over all candidates, best time of each:
                # clock drift during a sweep then favours no candidate
                times = {c: math.inf for c in cands}
                for _ in range(ROUNDS):
                    for c in cands:
                        times[c] = min(times[c], time_shape(ctx, elem, q, r, *c, reps))
                base = times[DEFAULT]
                (tb, mk), t_best = min(times.items(), key=lambda kv: kv[1])
                # keep the default unless a candidate is clearly faster
                if not t_best < (1.0 - MARGIN) * base:
                    (tb, mk), t_best = DEFAULT, base
                entries.append({"elem": elem, "q": q, "q_max": q_max, "r": r, "r_max": r_max,
                                "target_blocks": tb, "min_kspan": mk, "ms": t_best, "ms_default": base})
                log(f"[tune] {ELEM_NAME[elem]} q={q} r={r}: blocks {tb}, k span {mk}: "
                    f"{t_best:.4f} ms (default {base:.4f} ms)")
    return entries


ELEM_NAME = {2: "FP16 halves", 4: "FP32", 8: "FP64"}
ROW = re.compile(r"\{(\d+), (\d+)LL, (\d+), (\d+), (\d+)\},\s*// q (\d+), r (\d+): ([\d.]+) / ([\d.]+) ms")


def read_table(path: Path = TABLE) -> list[dict]:
    """The rows of an existing table, as tune() returns them."""
    out = []
    for m in ROW.finditer(path.read_text()):
        e, qm, rm, tb, mk, q, r, ms, msd = m.groups()
        out.append({"elem": int(e), "q": int(q), "q_max": int(qm), "r": int(r), "r_max": int(rm),
                    "target_blocks": int(tb), "min_kspan": int(mk), "ms": float(ms), "ms_default": float(msd)})
    return out


def render(entries: list[dict], device: str) -> str:
    lines = [
        "// Split-K shapes of k_dense_mfma (dense Q X), picked per (element size, q,",
        "// r) by cosmicblup_amd/tune.py from timings on an MI355X. Generated file:",
        "// rerun the tuner to replace it. An entry applies to q <= q_max and r <=",
        "// r_max (first match in this order); shapes past the table use 2048 blocks",
        "// and k spans of at least 256.",
        f"// Tuned on {device}, {time.strftime('%Y-%m-%d')}; ms = tuned / default time of one launch.",
        "#pragma once",
        "#include <cstdint>",
        "",
        "struct DenseTunedEntry {",
        "    int elem;            // bytes per element of Q and X (2: FP16 halves, 4: FP32, 8: FP64)",
        "    int64_t q_max;",
        "    int r_max;",
        "    int target_blocks;   // blocks aimed at (row tiles x k splits)",
        "    int min_kspan;       // smallest k range of one split",
        "};",
        "",
        "static const DenseTunedEntry kDenseTuned[] = {",
    ]
    for e in entries:
        lines.append(f"    {{{e['elem']}, {e['q_max']}LL, {e['r_max']}, {e['target_blocks']}, {e['min_kspan']}}},"
                     f"  // q {e['q']}, r {e['r']}: {e['ms']:.4f} / {e['ms_default']:.4f} ms")
    lines += ["};", ""]
    return "\n".join(lines)


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--dry-run", action="store_true", help="print the table instead of writing it")
    ap.add_argument("--out", type=Path, default=TABLE)
    ap.add_argument("--elems", type=int, nargs="+", choices=ELEMS, default=list(ELEMS),
                    help="element sizes to tune; the table's rows for the others are kept")
    args = ap.parse_args(argv)
    import torch

    ctx = _lib.Context(0)
    device = "device 0"
    if torch.cuda.is_available():
        props = torch.cuda.get_device_properties(0)
        device = f"{props.name} ({getattr(props, 'gcnArchName', 'unknown arch')})"
    new = tune(ctx, elems=tuple(args.elems))
    kept = [e for e in read_table() if e["elem"] not in args.elems] if TABLE.exists() else []
    order = {e: i for i, e in enumerate(ELEMS)}
    text = render(sorted(kept + new, key=lambda e: (order[e["elem"]], e["q"], e["r"])), device)
    if args.dry_run:
        print(text)
    else:
        args.out.write_text(text)
        print(f"[tune] wrote {args.out}")


if __name__ == "__main__":
    main()
