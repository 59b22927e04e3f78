"""Oracle (test infrastructure) for BED decode, SNP stats, integer Gram and
VanRaden G: ctypes over oracle/liborc.so (grm_oracle.c), plus the
eigen-bent inverse of src/core/matrix_io.cpp:488-501 in numpy/LAPACK.

Citations and pinning are in grm_oracle.c's header; invert_psd_matrix is
pinned by tests/test_oracle_grm.py (G_b G_b^-1 = I on a well-conditioned
matrix; floor of a singular centred G at eps).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liborc.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        vp, i64 = C.c_void_p, C.c_int64
        L.orc_snp_stats.argtypes = [vp, i64, i64, i64, C.c_float] + [vp] * 9
        L.orc_int_gram.argtypes = [vp, i64, i64, i64, vp, vp]
        L.orc_grm_vanraden.argtypes = [vp, i64, i64, i64, i64, C.c_float, vp, vp, vp]
        L.orc_grm_block_raw.argtypes = [vp, i64, i64, i64, vp, vp, i64, i64, i64, i64, i64, vp]
        L.orc_grm_rows_raw.argtypes = [vp, i64, i64, i64, vp, vp, vp, i64, i64, i64, i64, vp]
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _rows(bed: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(bed[3:])


def snp_stats(bed: np.ndarray, n: int, m: int, min_maf: float = 0.01) -> dict[str, np.ndarray]:
    return snp_stats_rows(_rows(bed), (n + 3) // 4, n, m, min_maf)


def snp_stats_rows(rows: np.ndarray, bps: int, n: int, m: int, min_maf: float = 0.01) -> dict[str, np.ndarray]:
    """snp_stats on SNP-major rows of `bps` bytes each (bps >= ceil(n/4))."""
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    out = {k: np.zeros(m, dtype=np.int64) for k in ("n_obs", "sum_x", "n_hom1", "n_het", "n_hom2")}
    out["valid"] = np.zeros(m, dtype=np.uint8)
    for k in ("freq_f", "mean_f", "twopq_f"):
        out[k] = np.zeros(m, dtype=np.float32)
    lib().orc_snp_stats(_p(rows), n, m, bps, C.c_float(min_maf), _p(out["n_obs"]), _p(out["sum_x"]),
                        _p(out["n_hom1"]), _p(out["n_het"]), _p(out["n_hom2"]), _p(out["valid"]),
                        _p(out["freq_f"]), _p(out["mean_f"]), _p(out["twopq_f"]))
    return out


def int_gram(bed: np.ndarray, n: int, m: int, valid: np.ndarray) -> np.ndarray:
    rows = _rows(bed)
    out = np.zeros((n, n), dtype=np.int64)
    v = np.ascontiguousarray(valid, dtype=np.uint8)
    lib().orc_int_gram(_p(rows), n, m, (n + 3) // 4, _p(v), _p(out))
    return out


def grm_vanraden(bed: np.ndarray, n: int, m: int, block_size: int = 2048,
                 min_maf: float = 0.01) -> tuple[np.ndarray, float, int]:
    rows = _rows(bed)
    G = np.zeros((n, n), dtype=np.float64)
    s = C.c_double()
    t = C.c_int64()
    lib().orc_grm_vanraden(_p(rows), n, m, (n + 3) // 4, block_size, C.c_float(min_maf), _p(G),
                           C.addressof(s), C.addressof(t))
    return G, s.value, t.value


def grm_block_raw(bed: np.ndarray, n: int, m: int, stats: dict, r0: int, nr: int, c0: int, nc: int,
                  block_size: int = 2048) -> np.ndarray:
    rows = _rows(bed)
    out = np.zeros((nr, nc), dtype=np.float64)
    lib().orc_grm_block_raw(_p(rows), n, m, (n + 3) // 4, _p(np.ascontiguousarray(stats["valid"])),
                            _p(np.ascontiguousarray(stats["mean_f"])), r0, nr, c0, nc, block_size,
                            _p(out))
    return out


def grm_rows_raw(rows: np.ndarray, nbytes: int, n: int, m: int, stats: dict, sel, c0: int = 0,
                 nc: int | None = None, block_size: int = 2048) -> np.ndarray:
    """Unnormalised Z Z^T of the listed samples `sel` against columns
    [c0, c0+nc), from SNP-major rows (m x nbytes uint8, nbytes >= ceil(n/4),
    e.g. a device BED copied to the host)."""
    nc = n - c0 if nc is None else nc
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    assert rows.shape[0] == m and rows.shape[1] == nbytes and nbytes >= (n + 3) // 4
    sel = np.ascontiguousarray(sel, dtype=np.int64)
    out = np.zeros((sel.size, nc), dtype=np.float64)
    lib().orc_grm_rows_raw(_p(rows), n, m, nbytes, _p(np.ascontiguousarray(stats["valid"])),
                           _p(np.ascontiguousarray(stats["mean_f"])), _p(sel), sel.size, c0, nc, block_size,
                           _p(out))
    return out


def num_threads() -> int:
    return lib().orc_num_threads()


def invert_psd_matrix(M: np.ndarray, ridge: float = 0.0) -> np.ndarray:
    """src/core/matrix_io.cpp:488-501: eigendecomposition, eigenvalues floored
    at eps = max(1e-12 * lambda_max, 1e-6), returns U diag(1/lambda) U^T."""
    M = np.array(M, dtype=np.float64, copy=True)
    if ridge > 0.0:
        M[np.diag_indices_from(M)] += ridge
    w, U = np.linalg.eigh(M)
    eps = max(1e-12 * w.max(), 1e-6)
    w = np.where(w < eps, eps, w)
    return (U / w) @ U.T


def bent_matrix(M: np.ndarray, ridge: float = 0.0) -> np.ndarray:
    """The matrix whose inverse invert_psd_matrix returns (U diag(lambda_floored) U^T);
    the dense-exact AI-REML re-inverts G^-1 to this (vce.cpp:2888-2895)."""
    M = np.array(M, dtype=np.float64, copy=True)
    if ridge > 0.0:
        M[np.diag_indices_from(M)] += ridge
    w, U = np.linalg.eigh(M)
    eps = max(1e-12 * w.max(), 1e-6)
    w = np.where(w < eps, eps, w)
    return (U * w) @ U.T
