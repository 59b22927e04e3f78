"""ctypes binding of libcosmicblup_amd.so (the C-ABI in include/cosmicblup_amd.h).

The library is loaded from this package directory only; a missing library
is an error, never a silent fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libcosmicblup_amd.so"

CB_OK = 0
CB_ERR_INVALID = -1
CB_ERR_OOM = -2
CB_ERR_DEVICE = -3
CB_ERR_NOT_CONVERGED = -4
CB_ERR_INTERNAL = -5
CB_ERR_IO = -6
CB_MATRIX_TXT = 0
CB_MATRIX_BIN = 1

CB_GRM_OUT_F64 = 1
CB_GRM_OUT_F32 = 2
CB_GRM_OUT_INT = 4

c_i64 = C.c_int64
c_vp = C.c_void_p
c_i64p = C.POINTER(C.c_int64)
c_dp = C.POINTER(C.c_double)


class GrmOptions(C.Structure):
    _fields_ = [("block_size", C.c_int64), ("min_maf", C.c_float), ("min_hwe", C.c_float),
                ("algorithm", C.c_int)]


class SolveOptions(C.Structure):
    _fields_ = [("precision", C.c_int), ("precond", C.c_int), ("stop_rule", C.c_int),
                ("batch_rhs", C.c_int), ("tol", C.c_double), ("max_iter", C.c_int64),
                ("stall_tol", C.c_double)]


class HinvOptions(C.Structure):
    _fields_ = [("tune_g", C.c_int), ("alpha", C.c_double), ("tau", C.c_double), ("omega", C.c_double),
                ("ridge", C.c_double)]


class MvOptions(C.Structure):
    _fields_ = [("max_iter", C.c_int64), ("tol", C.c_double)]


class RemlOptions(C.Structure):
    _fields_ = [("algorithm", C.c_int), ("max_iter", C.c_int64), ("tol", C.c_double),
                ("n_probes", C.c_int64), ("seed", C.c_uint64), ("trace_exact", C.c_int),
                ("solve", SolveOptions), ("block_cols", C.c_int64), ("mc_samples", C.c_int64),
                ("adaptive", C.c_int)]


CB_PREC_FP64 = 0
CB_PREC_MIXED = 1
CB_PREC_MIXED_F16 = 2
CB_PRECOND_JACOBI = 0
CB_PRECOND_BLOCK_JACOBI = 1
CB_STOP_RELATIVE = 0
CB_STOP_ABSOLUTE = 1
CB_VCE_AI = 0
CB_VCE_EM = 1
CB_VCE_AI_EXACT = 2
CB_VCE_AUTO = 3


class CosmicError(RuntimeError):
    """A non-zero status from the C-ABI (message from cb_last_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class NotConverged(CosmicError):
    pass


# name -> (restype, argtypes)
_SIGS = {
    "cb_context_create": (C.c_int, [C.c_int, C.POINTER(c_vp)]),
    "cb_context_destroy": (None, [c_vp]),
    "cb_last_error": (C.c_char_p, []),
    "cb_context_set_stream": (C.c_int, [c_vp, c_vp]),
    "cb_context_stream": (c_vp, [c_vp]),
    "cb_context_synchronize": (C.c_int, [c_vp]),
    "cb_context_set_profiling": (C.c_int, [c_vp, C.c_int]),
    "cb_context_set_graphs": (C.c_int, [c_vp, C.c_int]),
    "cb_grm_write": (C.c_int, [c_vp, C.c_char_p, C.c_int]),
    "cb_matrix_write": (C.c_int, [C.c_char_p, c_i64, c_vp, c_i64, C.c_int]),
    "cb_upper_packed_read": (C.c_int, [C.c_char_p, c_i64p, c_vp, c_i64]),
    "cb_context_reset_profile": (C.c_int, [c_vp]),
    "cb_context_kernel_time": (C.c_int, [c_vp, C.c_char_p, c_dp, c_i64p]),
    "cb_genotypes_from_bed": (C.c_int, [c_vp, c_vp, C.c_size_t, c_i64, c_i64, C.POINTER(c_vp)]),
    "cb_genotypes_from_device_rows": (C.c_int, [c_vp, c_vp, C.c_size_t, c_i64, c_i64, C.POINTER(c_vp)]),
    "cb_genotypes_destroy": (None, [c_vp]),
    "cb_genotypes_dims": (C.c_int, [c_vp, c_i64p, c_i64p]),
    "cb_grm_options_default": (None, [C.POINTER(GrmOptions)]),
    "cb_grm_create": (C.c_int, [c_vp, C.POINTER(GrmOptions), C.POINTER(c_vp)]),
    "cb_grm_destroy": (None, [c_vp]),
    "cb_grm_compute": (C.c_int, [c_vp, c_vp, C.c_int]),
    "cb_grm_copy_grm": (C.c_int, [c_vp, c_vp, c_i64]),
    "cb_grm_copy_grm_f32": (C.c_int, [c_vp, c_vp, c_i64]),
    "cb_grm_copy_int_gram": (C.c_int, [c_vp, c_vp, c_i64]),
    "cb_grm_device_ptr": (C.c_int, [c_vp, C.c_int, C.POINTER(c_vp), c_i64p]),
    "cb_grm_copy_grm_device": (C.c_int, [c_vp, c_vp, c_i64]),
    "cb_grm_sum2pq": (C.c_int, [c_vp, c_dp]),
    "cb_grm_total_snps": (C.c_int, [c_vp, c_i64p]),
    "cb_grm_num_samples": (C.c_int, [c_vp, c_i64p]),
    "cb_grm_copy_allele_freqs": (C.c_int, [c_vp, c_vp, c_i64]),
    "cb_grm_copy_snp_stats": (C.c_int, [c_vp, c_i64] + [c_vp] * 8),
    "cb_mme_create": (C.c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, C.POINTER(c_vp)]),
    "cb_mme_destroy": (None, [c_vp]),
    "cb_mme_dims": (C.c_int, [c_vp, c_i64p, c_i64p, c_i64p]),
    "cb_mme_set_qinv_dense": (C.c_int, [c_vp, c_vp, c_i64]),
    "cb_mme_set_qinv_dense_device": (C.c_int, [c_vp, c_vp, c_i64]),
    "cb_mme_set_qinv_csr": (C.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "cb_mme_rhs": (C.c_int, [c_vp, c_vp]),
    "cb_mme_apply": (C.c_int, [c_vp, C.c_double, c_i64, c_vp, c_vp]),
    "cb_solve_options_default": (None, [C.POINTER(SolveOptions)]),
    "cb_mme_solve": (C.c_int, [c_vp, C.c_double, C.POINTER(SolveOptions), c_i64, c_vp, c_vp, c_i64p, c_dp]),
    "cb_mme_solve_device": (C.c_int, [c_vp, C.c_double, C.POINTER(SolveOptions), c_i64, c_vp, c_i64, c_vp, c_i64,
                                      c_i64p, c_dp]),
    "cb_tune_dense_time": (C.c_int, [c_vp, C.c_int, c_i64, C.c_int, C.c_int, C.c_int, C.c_int, c_dp]),
    "cb_tune_dense_lookup": (C.c_int, [C.c_int, c_i64, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cb_mme_solve_device_noalloc": (C.c_int, [c_vp, C.c_double, C.POINTER(SolveOptions), c_i64, c_vp, c_i64,
                                              c_vp, c_i64, c_i64p, c_dp]),
    "cb_mme_beta_cov": (C.c_int, [c_vp, C.c_double, C.c_double, C.POINTER(SolveOptions), c_vp, c_vp, c_i64p]),
    "cb_reml_options_default": (None, [C.POINTER(RemlOptions)]),
    "cb_reml_run": (C.c_int, [c_vp, C.POINTER(RemlOptions), C.c_double, C.c_double, c_dp, c_dp, c_i64p,
                              C.POINTER(C.c_int), c_vp, c_vp, c_vp, c_vp, c_vp]),
    "cb_reml_logl": (C.c_int, [c_vp, C.c_double, C.c_double, c_dp, c_dp]),
    "cb_reml_score": (C.c_int, [c_vp, C.POINTER(RemlOptions), C.c_double, C.c_double, c_vp, c_vp, c_dp, c_vp]),
    "cb_reml_copy_probes": (C.c_int, [c_vp, C.POINTER(RemlOptions), c_vp, c_i64]),
    "cb_reml_branch": (C.c_int, [c_vp, C.POINTER(RemlOptions), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int), c_i64p, c_i64p]),
    "cb_invert_psd": (C.c_int, [c_vp, c_vp, c_i64, c_i64, C.c_double, c_vp, c_i64, c_dp, c_dp]),
    "cb_invert_psd_device": (C.c_int, [c_vp, c_vp, c_i64, c_i64, C.c_double, c_vp, c_i64, c_dp, c_dp]),
    "cb_mme_set_qinv_split": (C.c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_i64]),
    "cb_hinv_options_default": (None, [C.POINTER(HinvOptions)]),
    "cb_hinv_blocks": (C.c_int, [c_vp, c_i64, c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, C.POINTER(HinvOptions), c_vp,
                                 c_i64, c_vp, c_i64, c_vp]),
    "cb_mv_options_default": (None, [C.POINTER(MvOptions)]),
    "cb_mv_vce": (C.c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_i64, C.POINTER(MvOptions), c_vp, c_vp,
                            c_i64p, C.POINTER(C.c_int)]),
    "cb_mv_solve": (C.c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "cb_ainv_henderson": (C.c_int, [c_i64, c_vp, c_vp, C.POINTER(c_vp)]),
    "cb_ainv_nnz": (C.c_int, [c_vp, c_i64p]),
    "cb_ainv_copy_csr": (C.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "cb_ainv_copy_inbreeding": (C.c_int, [c_vp, c_vp, c_vp]),
    "cb_ainv_destroy": (None, [c_vp]),
}

_lib = None


def exported_names() -> list[str]:
    return list(_SIGS)


def lib() -> C.CDLL:
    """Load the in-tree library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python cosmicblup_amd/build.py` (there is no CPU fallback)")
    L = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int) -> None:
    if status != CB_OK:
        msg = lib().cb_last_error().decode(errors="replace")
        if status == CB_ERR_NOT_CONVERGED:
            raise NotConverged(status, msg)
        raise CosmicError(status, msg)


def ptr(a) -> int:
    """Address of a numpy array's data (C-contiguous required)."""
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


class Context:
    """Owns a cb_context (device + HIP stream + kernel timing)."""

    def __init__(self, device: int = 0):
        h = c_vp()
        check(lib().cb_context_create(device, C.byref(h)))
        self.handle = h
        self.device = device

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().cb_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, hip_stream: int | None) -> None:
        check(lib().cb_context_set_stream(self.handle, hip_stream))

    @property
    def stream(self) -> int:
        return lib().cb_context_stream(self.handle)

    def synchronize(self) -> None:
        check(lib().cb_context_synchronize(self.handle))

    def set_profiling(self, on: bool) -> None:
        check(lib().cb_context_set_profiling(self.handle, int(on)))

    def set_graphs(self, on: bool) -> None:
        """Replay PCG iterations as HIP graphs, or launch them one by one (default)."""
        check(lib().cb_context_set_graphs(self.handle, int(on)))

    def reset_profile(self) -> None:
        check(lib().cb_context_reset_profile(self.handle))

    def kernel_time(self, name: str) -> tuple[float, int]:
        ms = C.c_double()
        cnt = C.c_int64()
        check(lib().cb_context_kernel_time(self.handle, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx
