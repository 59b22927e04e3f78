"""Fixed-effect design — host mirror of buildFixedDesignGeneric
(src/core/design.cpp:13-109).

Column 0 is the mean; each categorical factor contributes (levels - 1)
dummy columns with levels in string (std::set) order and the first level
dropped; numeric covariates follow as one column each. Rows are CSR.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class FixedDesignG:
    p: int
    names: list[str]
    ptr: np.ndarray   # int64, n+1
    col: np.ndarray   # int32
    val: np.ndarray   # float64
    factor_cols: dict[str, list[int]] = field(default_factory=dict)


def build_fixed_design(cats: list[list[str]], nums: list[list[float]] | None = None,
                       cat_names: list[str] | None = None,
                       num_names: list[str] | None = None) -> FixedDesignG:
    """cats[r][j]: level string of factor j for record r ("" = missing level);
    nums[r][j]: covariate values."""
    n = len(cats)
    ncat = len(cats[0]) if n else 0
    levels = []
    for j in range(ncat):
        lv = sorted({r[j] for r in cats if r[j] != ""})  # std::set<std::string> order
        levels.append(lv)
    names = ["mu"]
    factor_cols = {"mu": [0]}
    p = 1
    for j, lv in enumerate(levels):
        if not lv:
            continue
        prefix = cat_names[j] if cat_names and j < len(cat_names) else f"factor{j + 1}"
        for lvl in lv[1:]:
            names.append(f"{prefix}_{lvl}")
            factor_cols.setdefault(prefix, []).append(p)
            p += 1
    nq = len(nums[0]) if nums else 0
    for j in range(nq):
        name = num_names[j] if num_names and j < len(num_names) else f"cov{j + 1}"
        names.append(name)
        factor_cols.setdefault(name, []).append(p)
        p += 1
    # column of each (factor, level), base level -> none
    colmap = []
    c = 1
    for lv in levels:
        m = {}
        for lvl in lv[1:]:
            m[lvl] = c
            c += 1
        colmap.append(m)
    ptr = [0]
    col: list[int] = []
    val: list[float] = []
    for r in range(n):
        col.append(0)
        val.append(1.0)
        for j in range(ncat):
            cc = colmap[j].get(cats[r][j])
            if cc is not None:
                col.append(cc)
                val.append(1.0)
        for j in range(nq):
            col.append(c + j)
            val.append(float(nums[r][j]))
        ptr.append(len(col))
    return FixedDesignG(p=p, names=names, ptr=np.asarray(ptr, dtype=np.int64),
                        col=np.asarray(col, dtype=np.int32), val=np.asarray(val, dtype=np.float64),
                        factor_cols=factor_cols)


def dense_x(fd: FixedDesignG) -> np.ndarray:
    n = fd.ptr.size - 1
    X = np.zeros((n, fd.p))
    for r in range(n):
        for k in range(fd.ptr[r], fd.ptr[r + 1]):
            X[r, fd.col[k]] += fd.val[k]
    return X
