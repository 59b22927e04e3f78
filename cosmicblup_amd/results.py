"""Result files of a single-trait run with one random effect — host mirror
of GSOutput::writeResults (src/solve/result_writer.cpp:64-433): .vars
(variances, h2 with delta-method SE and a normal-approximation p value),
.beta (fixed effects) and .rand (random effects averaged over each
animal's records, mean residual, record count).

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
