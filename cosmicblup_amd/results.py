"""Result files — host mirror of GSOutput::writeResults
(src/solve/result_writer.cpp:64-562) for a single-trait run with one random
effect: .vars (variances, h2 with delta-method SE and a normal-approximation
p value), .beta (fixed effects with SEs), .rand (random effects averaged
over each animal's records, mean residual, record count) and .obsrand (per
record: fixed part, random part, fitted value, residual); and of the
multi-trait outputs: .mv.vars (variance_estimator.cpp:303-311) and
GSOutput::writeMultiTrait's .beta / .rand (result_writer.cpp:672-761).

Numbers are written as a C++ stream writes a double by default (%g, six
significant digits)."""
from __future__ import annotations

import math

import numpy as np

from .design import FixedDesignG

MODEL_TAGS = {"gblup": "GA", "pblup": "PA", "ssgblup": "HA"}


def fmt(v: float) -> str:
    return format(float(v), "g")


def calc_pval(val: float, se: float) -> float:
    """result_writer.cpp:20-24: erfc(z / sqrt 2) with z = val / se, 1 when
    se < 1e-12."""
    if se < 1e-12:
        return 1.0
    return math.erfc(val / se / math.sqrt(2.0))


def format_pval(p: float) -> str:
    return "<2e-16" if p < 2e-16 else fmt(p)


def _h2_se(k: int, vars_all: list[float], ses_all: list[float]) -> tuple[float, float]:
    """Share of component k in the total and its delta-method SE, the other
    components (the residual included) entering through d h / d other."""
    total = sum(vars_all)
    h2 = vars_all[k] / total if total > 1e-9 else 0.0
    denom = total * total
    d_self = (total - vars_all[k]) / denom
    d_other = -vars_all[k] / denom
    var = d_self * d_self * ses_all[k] * ses_all[k]
    for j, s in enumerate(ses_all):
        if j != k:
            var += d_other * d_other * s * s
    return h2, math.sqrt(max(0.0, var))


def write_vars(out_prefix: str, model: str, var_u: float, var_e: float, se_u: float = 0.0,
               se_e: float = 0.0) -> str:
    """The .vars table (result_writer.cpp:76-205) for one random effect."""
    tag = MODEL_TAGS.get(model, "PA")
    path = out_prefix + ".vars"
    # the random effect's h2 adds its own SE term and then the residual's
    h2, se_h2 = _h2_se(0, [var_u, var_e], [se_u, se_e])
    h2e, se_h2e = _h2_se(1, [var_u, var_e], [se_u, se_e])
    with open(path, "w") as f:
        f.write("Item\tVar\tVar_SE\th2\th2_SE\th2_Pr(Chisq)\n")
        f.write(f"{tag}\t{fmt(var_u)}\t{fmt(se_u)}\t{fmt(h2)}\t{fmt(se_h2)}\t{format_pval(calc_pval(h2, se_h2))}\n")
        f.write(f"e\t{fmt(var_e)}\t{fmt(se_e)}\t{fmt(h2e)}\t{fmt(se_h2e)}\t{format_pval(calc_pval(h2e, se_h2e))}\n")
    return path


def write_beta(out_prefix: str, fd: FixedDesignG, beta, beta_se=None) -> str:
    """The .beta table (result_writer.cpp:207-241): level names, estimates,
    SE or nan."""
    path = out_prefix + ".beta"
    beta = np.asarray(beta, dtype=np.float64)
    with open(path, "w") as f:
        f.write("Levels\tEstimation\tSE\n")
        for i in range(fd.p):
            se = fmt(beta_se[i]) if beta_se is not None and i < len(beta_se) else "nan"
            f.write(f"{fd.names[i]}\t{fmt(beta[i])}\t{se}\n")
    return path


def write_rand(out_prefix: str, model: str, ids: list[str], rec_aid, y, fd: FixedDesignG, beta, u) -> str:
    """The .rand table (result_writer.cpp:296-433) for one random effect
    indexed by the record's 1-based aid: per animal the mean over its
    records of u, the mean residual y - (x'b + u) and the record count;
    animals with records first (in order of their first record), then the
    rest sorted by id with their u, nan residual and count 0. Sums run in
    record order, as the reference's."""
    n_ind = len(ids)
    rec_aid = np.asarray(rec_aid, dtype=np.int64)
    beta = np.asarray(beta, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    res_sum = [0.0] * n_ind
    comp_sum = [0.0] * n_ind
    cnt = [0] * n_ind
    for i in range(rec_aid.size):
        xb = 0.0
        for k in range(fd.ptr[i], fd.ptr[i + 1]):
            c = int(fd.col[k])
            if 0 <= c < fd.p:
                xb += float(fd.val[k]) * float(beta[c])
        a = int(rec_aid[i]) - 1
        val = float(u[a]) if 0 <= a < u.size else 0.0
        resid = float(y[i]) - (xb + val)
        if 0 <= a < n_ind:
            res_sum[a] += resid
            cnt[a] += 1
            comp_sum[a] += val
    order, seen = [], [False] * n_ind
    for a in rec_aid:
        a = int(a) - 1
        if 0 <= a < n_ind and not seen[a]:
            seen[a] = True
            order.append(a)
    order += sorted((i for i in range(n_ind) if not seen[i]), key=lambda i: ids[i])
    path = out_prefix + ".rand"
    col = MODEL_TAGS.get(model, "PA")
    with open(path, "w") as f:
        f.write(f"ID\t{col}\tresiduals\tn_obs\n")
        for i in order:
            name = ids[i] if ids[i] else f"UNK_{i + 1}"
            if cnt[i] > 0:
                f.write(f"{name}\t{fmt(comp_sum[i] / cnt[i])}\t{fmt(res_sum[i] / cnt[i])}\t{cnt[i]}\n")
            else:
                g = fmt(u[i]) if i < u.size else "nan"
                f.write(f"{name}\t{g}\tnan\t0\n")
    return path


def write_obsrand(out_prefix: str, model: str, rec_ids: list[str], y, fd: FixedDesignG, beta, u, rec_aid) -> str:
    """The .obsrand table (result_writer.cpp:345-375 values, 548-562 file):
    per record (1-based), its id, y, x'b, the random effect u[aid - 1] (0
    for a record without a level), fitted = x'b + u and y - fitted."""
    rec_aid = np.asarray(rec_aid, dtype=np.int64)
    beta = np.asarray(beta, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    col = MODEL_TAGS.get(model, "PA")
    path = out_prefix + ".obsrand"
    with open(path, "w") as f:
        f.write(f"Record\tID\tY\tFixed\t{col}\tFitted\tResidual\n")
        for i in range(rec_aid.size):
            xb = 0.0
            for k in range(fd.ptr[i], fd.ptr[i + 1]):
                c = int(fd.col[k])
                if 0 <= c < fd.p:
                    xb += float(fd.val[k]) * float(beta[c])
            a = int(rec_aid[i]) - 1
            val = float(u[a]) if 0 <= a < u.size else 0.0
            fitted = xb + val
            f.write(f"{i + 1}\t{rec_ids[i]}\t{fmt(y[i])}\t{fmt(xb)}\t{fmt(val)}\t{fmt(fitted)}\t"
                    f"{fmt(float(y[i]) - fitted)}\n")
    return path


def write_mv_vars(out_prefix: str, Vg, Ve) -> str:
    """The .mv.vars table of the multi-trait VCE (variance_estimator.cpp:
    303-311): one line per (trait i, trait j), 1-based, with Vg and Ve."""
    Vg, Ve = np.asarray(Vg, dtype=np.float64), np.asarray(Ve, dtype=np.float64)
    path = out_prefix + ".mv.vars"
    with open(path, "w") as f:
        f.write("Trait1\tTrait2\tVg\tVe\n")
        for i in range(Vg.shape[0]):
            for j in range(Vg.shape[1]):
                f.write(f"{i + 1}\t{j + 1}\t{fmt(Vg[i, j])}\t{fmt(Ve[i, j])}\n")
    return path


def write_multi_trait(out_prefix: str, beta, u, ids: list[str], trait_names: list[str] | None = None,
                      effect_names: list[str] | None = None) -> tuple[str, str]:
    """GSOutput::writeMultiTrait (result_writer.cpp:672-761): .beta with one
    row per fixed effect and one column per trait, .rand with one row per
    individual (EBV per trait). beta: p x t, u: n x t. Trait names default to
    Trait_j; effect names to Intercept (p = 1) or Covar_k when no names of
    length p are given."""
    beta = np.asarray(beta, dtype=np.float64).reshape(-1, np.asarray(u).shape[1])
    u = np.asarray(u, dtype=np.float64)
    p, t = beta.shape
    names = list(trait_names) if trait_names is not None and len(trait_names) >= t else \
        [f"Trait_{j + 1}" for j in range(t)]
    if effect_names is not None and len(effect_names) == p:
        eff = list(effect_names)
    elif p == 1:
        eff = ["Intercept"]
    else:
        eff = [f"Covar_{k + 1}" for k in range(p)]
    bpath, rpath = out_prefix + ".beta", out_prefix + ".rand"
    with open(bpath, "w") as f:
        f.write("Effect" + "".join(f"\t{names[j]}" for j in range(t)) + "\n")
        for k in range(p):
            f.write(eff[k] + "".join(f"\t{fmt(beta[k, j])}" for j in range(t)) + "\n")
    with open(rpath, "w") as f:
        f.write("ID" + "".join(f"\t{names[j]}" for j in range(t)) + "\n")
        for i in range(u.shape[0]):
            f.write(ids[i] + "".join(f"\t{fmt(u[i, j])}" for j in range(t)) + "\n")
    return bpath, rpath
