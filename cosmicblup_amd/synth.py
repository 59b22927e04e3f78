"""Seeded synthetic inputs in the reference's file formats.

Genotypes: per-SNP allele frequency p_j ~ U(lo, hi) of the counted (.bim
column 5) allele, dosages ~ Binomial(2, p_j) under HWE, MCAR missing calls.
BED codes (plink_reader.cpp:17-20): dosage 2 -> 00, 1 -> 10, 0 -> 11,
missing -> 01; sample i at byte i>>2, bits 2(i&3); magic 6C 1B 01.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

BED_MAGIC = np.array([0x6C, 0x1B, 0x01], dtype=np.uint8)
# dosage (0,1,2,missing=-1 -> index 3) -> 2-bit code
_CODE = np.array([3, 2, 0, 1], dtype=np.uint8)


def simulate_dosages(n: int, m: int, maf=(0.05, 0.5), missing_rate: float = 0.0,
                     seed: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """Return (dosage int8 n x m with -1 = missing, p float64 m)."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(maf[0], maf[1], size=m)
    x = rng.binomial(2, p, size=(n, m)).astype(np.int8)
    if missing_rate > 0:
        x[rng.random((n, m)) < missing_rate] = -1
    return x, p


def encode_bed(dosage: np.ndarray) -> np.ndarray:
    """SNP-major .bed image (with magic) of an n x m int8 dosage matrix."""
    n, m = dosage.shape
    bps = (n + 3) // 4
    idx = np.where(dosage < 0, 3, dosage).astype(np.int64)
    codes = _CODE[idx].T  # m x n
    pad = bps * 4 - n
    if pad:
        codes = np.concatenate([codes, np.zeros((m, pad), dtype=np.uint8)], axis=1)
    c = codes.reshape(m, bps, 4).astype(np.uint8)
    packed = c[:, :, 0] | (c[:, :, 1] << 2) | (c[:, :, 2] << 4) | (c[:, :, 3] << 6)
    return np.concatenate([BED_MAGIC, packed.reshape(-1)])


def decode_bed(bed: np.ndarray, n: int, m: int) -> np.ndarray:
    """Inverse of encode_bed: n x m int8 dosages, -1 for missing."""
    bps = (n + 3) // 4
    if bed[0] != 0x6C or bed[1] != 0x1B or bed[2] != 0x01:
        raise RuntimeError("Invalid .bed file magic number")
    rows = bed[3:3 + m * bps].reshape(m, bps)
    codes = np.stack([(rows >> (2 * k)) & 3 for k in range(4)], axis=2).reshape(m, bps * 4)[:, :n]
    table = np.array([2, -1, 1, 0], dtype=np.int8)
    return table[codes].T.copy()


def write_plink(prefix: str, dosage: np.ndarray, iids=None) -> None:
    n, m = dosage.shape
    prefix = str(prefix)
    Path(prefix).parent.mkdir(parents=True, exist_ok=True)
    encode_bed(dosage).tofile(prefix + ".bed")
    iids = iids or [f"id{i + 1:07d}" for i in range(n)]
    with open(prefix + ".fam", "w") as f:
        for iid in iids:
            f.write(f"{iid} {iid} 0 0 0 -9\n")
    with open(prefix + ".bim", "w") as f:
        for j in range(m):
            f.write(f"1 snp{j + 1} 0 {j + 1} A B\n")


def bed_on_device(n: int, m: int, maf=(0.01, 0.5), missing_rate: float = 0.0, seed: int = 1,
                  device: str = "cuda"):
    """Large synthetic BED generated directly in device memory with torch.

    Returns (rows uint8 tensor m x row_pitch, row_pitch) once the device has
    finished writing them (the library reads them on its own stream). Same
    distribution as simulate_dosages; the random stream is torch's, so it is
    not the numpy one.
    """
    import torch

    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    bps = (n + 3) // 4
    pitch = (bps + 15) // 16 * 16
    rows = torch.zeros((m, pitch), dtype=torch.uint8, device=device)
    p = torch.empty(m, dtype=torch.float32, device=device).uniform_(maf[0], maf[1], generator=gen)
    chunk = max(1, (1 << 28) // max(n, 1))
    for j0 in range(0, m, chunk):
        j1 = min(m, j0 + chunk)
        pj = p[j0:j1, None]
        u = torch.rand((j1 - j0, bps * 4), device=device, generator=gen)
        # dosage 2 with prob p^2, 1 with 2p(1-p), 0 otherwise
        code = torch.full_like(u, 3, dtype=torch.uint8)
        code[u < pj * pj + 2 * pj * (1 - pj)] = 2
        code[u < pj * pj] = 0
        if missing_rate > 0:
            code[torch.rand(u.shape, device=device, generator=gen) < missing_rate] = 1
        code[:, n:] = 0
        c = code.view(j1 - j0, bps, 4)
        packed = c[:, :, 0] | (c[:, :, 1] << 2) | (c[:, :, 2] << 4) | (c[:, :, 3] << 6)
        rows[j0:j1, :bps] = packed
        del u, code, c, packed
    torch.cuda.synchronize(device)
    return rows, pitch


def simulate_phenotype(dosage: np.ndarray, n_qtl: int = 500, h2: float = 0.3, n_groups: int = 20,
                       seed: int = 2) -> tuple[np.ndarray, list[str]]:
    """y = mu + CG effect + g + e: g from n_qtl QTL with N(0,1) effects scaled
    to variance h2, e ~ N(0, 1 - h2), contemporary groups ~ N(0, 1) assigned
    uniformly. Group labels are zero-padded so string order is numeric order."""
    rng = np.random.default_rng(seed)
    n, m = dosage.shape
    qtl = rng.choice(m, size=min(n_qtl, m), replace=False)
    eff = rng.standard_normal(qtl.size)
    xq = np.where(dosage[:, qtl] < 0, 0, dosage[:, qtl]).astype(np.float64)
    g = (xq - xq.mean(axis=0)) @ eff
    g *= np.sqrt(h2 / max(g.var(), 1e-12))
    e = rng.standard_normal(n) * np.sqrt(1.0 - h2)
    grp = rng.integers(0, n_groups, size=n)
    ge = rng.standard_normal(n_groups)
    y = 10.0 + ge[grp] + g + e
    width = len(str(n_groups))
    labels = [f"cg{int(c):0{width}d}" for c in grp]
    return y, labels


def simulate_pedigree(n_gen: int, per_gen: int, unknown: float = 0.05, seed: int = 3):
    """Discrete generations of per_gen animals; sire and dam of every
    non-founder drawn uniformly from the previous generation, each unknown
    with probability `unknown`. Returns (dam, sire) as int32 row indices
    (-1 unknown); parents precede offspring."""
    rng = np.random.default_rng(seed)
    n = n_gen * per_gen
    dam = np.full(n, -1, dtype=np.int32)
    sire = np.full(n, -1, dtype=np.int32)
    for g in range(1, n_gen):
        lo = (g - 1) * per_gen
        s = lo + rng.integers(0, per_gen, per_gen)
        d = lo + rng.integers(0, per_gen, per_gen)
        s[rng.random(per_gen) < unknown] = -1
        d[rng.random(per_gen) < unknown] = -1
        sire[g * per_gen:(g + 1) * per_gen] = s
        dam[g * per_gen:(g + 1) * per_gen] = d
    return dam, sire


def simulate_pedigree_phenotype(dam, sire, dii, n_rec: int, h2: float = 0.25, n_groups: int = 2000,
                                seed: int = 4):
    """Breeding values down the pedigree, u_i = (u_s + u_d)/2 + sqrt(d_ii)
    N(0, h2) (unknown parent = 0), then n_rec distinct animals phenotyped:
    y = 10 + CG + u + e, e ~ N(0, 1 - h2), CG ~ N(0, 1) over n_groups groups.
    Returns (record animal rows int32, y, group index per record)."""
    rng = np.random.default_rng(seed)
    n = dam.size
    u = np.zeros(n)
    ms = rng.standard_normal(n) * np.sqrt(np.asarray(dii) * h2)
    # parents precede offspring: one pass in index order, vectorised per block
    # of animals whose parents all lie before the block
    i = 0
    while i < n:
        j = i + 1
        lim = min(n, i + 65536)
        while j < lim and max(dam[j], sire[j]) < i:
            j += 1
        pu = np.where(sire[i:j] >= 0, u[np.maximum(sire[i:j], 0)], 0.0)
        pd = np.where(dam[i:j] >= 0, u[np.maximum(dam[i:j], 0)], 0.0)
        u[i:j] = 0.5 * (pu + pd) + ms[i:j]
        i = j
    rec = np.sort(rng.choice(n, n_rec, replace=False)).astype(np.int32)
    grp = rng.integers(0, n_groups, n_rec)
    y = 10.0 + rng.standard_normal(n_groups)[grp] + u[rec] + rng.standard_normal(n_rec) * np.sqrt(1.0 - h2)
    return rec, y, grp
