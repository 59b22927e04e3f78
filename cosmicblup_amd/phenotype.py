"""Phenotype and variance files — host mirror of readPhenGeneric
(src/io/phenotype_loader.cpp:18-164) with its string helpers
(src/core/string_utils.h:16-42), and readVars (src/core/matrix_io.cpp:
416-465).

Records come out in file order with the 1-based numeric id of the
relationship matrix (idmap) and feed build_fixed_design (design.py)."""
from __future__ import annotations

import re
from dataclasses import dataclass, field

_WS = " \t\r\n"
_STOD = re.compile(r"[ \t\n\v\f\r]*[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan)",
                   re.IGNORECASE)


def trim_copy(s: str) -> str:
    return s.strip(_WS)


def detect_delim(first_line: str) -> str:
    return "," if "," in first_line else " "


def is_integer_string(s: str) -> bool:
    return bool(s) and all(c.isdigit() and c.isascii() for c in s)


def is_missing_token(s: str) -> bool:
    t = trim_copy(s)
    return t == "" or t.lower() in ("na", "nan", ".", "null")


def stod(s: str) -> float:
    """std::stod: the longest leading decimal (or inf / nan) prefix after
    leading whitespace; nothing to convert raises ValueError."""
    m = _STOD.match(s)
    if not m:
        raise ValueError(f"stod: no conversion: {s!r}")
    return float(m.group(0))


@dataclass
class GenRecord:
    idstr: str
    aid: int          # 1-based id in the relationship matrix
    y: float
    cats: list[str] = field(default_factory=list)
    nums: list[float] = field(default_factory=list)
    rand_cats: list[str] = field(default_factory=list)


@dataclass
class PhenStats:
    n_rows_total: int = 0
    n_dropped_missing: int = 0
    n_not_in_id: int = 0
    n_used: int = 0


def _split(row: str, delim: str) -> list[str]:
    if delim == " ":
        return row.split()
    return [trim_copy(t) for t in row.split(",")]


def read_phen(phen_file: str, idmap: dict[str, int], trait_pos_1based: int = 2,
              dcovar_pos_1based: list[int] | None = None, qcovar_pos_1based: list[int] | None = None,
              rand_pos_1based: list[int] | None = None, dcovar_names: list[str] | None = None,
              qcovar_names: list[str] | None = None, rand_names: list[str] | None = None,
              trait_name: str = "", id_col_name: str = "") -> tuple[list[GenRecord], PhenStats]:
    """readPhenGeneric. Columns are found by name (exact, then case-insensitive,
    or an all-digit string taken as a 0-based index) or by 1-based position.
    The name lists are filled in place from the header when positions are
    used, as the reference does. A row is dropped (n_dropped_missing) when it
    is short, has an empty id, a missing or unparsable trait, or a missing
    covariate; n_not_in_id counts ids absent from idmap."""
    dcovar_names = dcovar_names if dcovar_names is not None else []
    qcovar_names = qcovar_names if qcovar_names is not None else []
    rand_names = rand_names if rand_names is not None else []
    try:
        f = open(phen_file)
    except OSError:
        raise RuntimeError("Cannot open phenotype file: " + phen_file)
    with f:
        lines = f.read().split("\n")
    header = lines[0] if lines else ""
    delim = detect_delim(header)
    headers: list[str] = []
    if header[:1] in (" ", "\t"):
        headers.append("ID")
    headers += header.split() if delim == " " else [trim_copy(t) for t in header.split(",")]

    def find_col_index(name: str) -> int:
        if is_integer_string(name):
            return int(name)
        for i, h in enumerate(headers):
            if h == name or h.lower() == name.lower():
                return i
        return -1

    id_idx = find_col_index(id_col_name) if id_col_name else -1
    if id_idx < 0:
        for cand in ("id", "ID", "Id", "IID"):
            id_idx = find_col_index(cand)
            if id_idx >= 0:
                break
    if id_idx < 0:
        raise RuntimeError("Phenotype file missing required ID column")
    trait_idx = find_col_index(trait_name) if trait_name else -1
    if trait_idx < 0:
        trait_idx = trait_pos_1based - 1

    def resolve(names: list[str], positions: list[int] | None, what: str, fallback: str) -> list[int]:
        cols = []
        if names:
            for nm in names:
                idx = find_col_index(nm)
                if idx < 0:
                    raise RuntimeError(f"Cannot find {what} column: {nm}")
                cols.append(idx)
        else:
            for v in positions or []:
                cols.append(v - 1)
                names.append(headers[v - 1] if 0 <= v - 1 < len(headers) else f"{fallback}{v}")
        return cols

    dcols = resolve(dcovar_names, dcovar_pos_1based, "discrete covariate", "factor")
    qcols = resolve(qcovar_names, qcovar_pos_1based, "quantitative covariate", "cov")
    rcols = resolve(rand_names, rand_pos_1based, "random effect", "rand")

    recs: list[GenRecord] = []
    st = PhenStats()
    for line in lines[1:]:
        row = trim_copy(line)
        if not row:
            continue
        st.n_rows_total += 1
        vals = _split(row, delim)
        if trait_idx < 0:
            continue
        need = max([id_idx, trait_idx] + dcols + qcols + rcols)
        if len(vals) <= need:
            st.n_dropped_missing += 1
            continue
        idstr = vals[id_idx]
        if not idstr:
            st.n_dropped_missing += 1
            continue
        if idstr not in idmap:
            st.n_not_in_id += 1
            continue
        ystr = vals[trait_idx]
        if is_missing_token(ystr):
            st.n_dropped_missing += 1
            continue
        try:
            yv = stod(ystr)
        except ValueError:
            st.n_dropped_missing += 1
            continue
        r = GenRecord(idstr, idmap[idstr], yv)
        keep = True
        for dc in dcols:
            if is_missing_token(vals[dc]):
                keep = False
                break
            r.cats.append(vals[dc])
        if keep:
            for rc in rcols:
                if is_missing_token(vals[rc]):
                    keep = False
                    break
                r.rand_cats.append(vals[rc])
        if keep:
            for qc in qcols:
                if is_missing_token(vals[qc]):
                    keep = False
                    break
                try:
                    r.nums.append(stod(vals[qc]))
                except ValueError:
                    keep = False
                    break
        if not keep:
            st.n_dropped_missing += 1
            continue
        recs.append(r)
    st.n_used = len(recs)
    return recs, st


def read_vars(filename: str) -> tuple[float, float]:
    """readVars: "key value" lines; a first key of item / var (any case) is a
    header; e / error / residual is the residual variance, the first other
    key the random-effect variance. Falls back to the first two pairs when
    no line parses; both zero raises."""
    try:
        f = open(filename)
    except OSError:
        raise RuntimeError("Cannot open variance file: " + filename)
    with f:
        text = f.read()
    vu = ve = 0.0
    found_u = found_e = False
    for line in text.split("\n"):
        s = trim_copy(line)
        if not s:
            continue
        toks = s.split()
        key = toks[0]
        klow = key.lower()
        if klow in ("item", "var"):
            continue
        if len(toks) < 2:
            continue
        try:
            val = stod(toks[1])
        except ValueError:
            continue
        if klow in ("e", "error", "residual"):
            ve, found_e = val, True
        elif not found_u:
            vu, found_u = val, True
    if not found_u and not found_e:
        toks = text.split()
        pairs = []
        for k in range(0, len(toks) - 1, 2):
            try:
                pairs.append((toks[k], stod(toks[k + 1])))
            except ValueError:
                break
            if len(pairs) == 2:
                break
        for i, (k, v) in enumerate(pairs):
            if k.lower() == "e":
                ve, found_e = v, True
            elif i == 0 or not found_u:
                vu, found_u = v, True
    if vu == 0 and ve == 0:
        raise RuntimeError("Cannot parse variance components from: " + filename)
    return vu, ve
