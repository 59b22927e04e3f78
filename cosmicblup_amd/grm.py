"""GrmBuilder — host mirror of src/core/grm_builder.{h,cpp} (PLINK, VanRaden-1).

Same names and meaning as the reference: Options, compute(reader),
getGrm(), getSum2pq(), getTotalSnps(), getAlleleFreqs(). The work runs in
libcosmicblup_amd.so (cb_grm_*); G stays on the device until copied.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .plink import PlinkReader


@dataclass
class Options:
    """grm_builder.h:14-29 (fields used by the PLINK VanRaden path)."""
    thread_num: int = 1
    block_size: int = 2048
    min_maf: float = 0.01
    min_hwe: float = 0.0
    algorithm: int = 1


class GrmBuilder:
    def __init__(self, options: Options | None = None, ctx: _lib.Context | None = None):
        self.options = options or Options()
        self.ctx = ctx or _lib.default_context()
        o = _lib.GrmOptions()
        _lib.lib().cb_grm_options_default(C.byref(o))
        o.block_size = self.options.block_size
        o.min_maf = self.options.min_maf
        o.min_hwe = self.options.min_hwe
        o.algorithm = self.options.algorithm
        h = C.c_void_p()
        _lib.check(_lib.lib().cb_grm_create(self.ctx.handle, C.byref(o), C.byref(h)))
        self.handle = h
        self._n = 0
        self._m = 0

    def close(self) -> None:
        if getattr(self, "handle", None):
            _lib.lib().cb_grm_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # grm_builder.cpp:63-405
    def compute(self, reader: PlinkReader, keep_indices=None,
                outputs: int = _lib.CB_GRM_OUT_F64) -> None:
        if keep_indices:
            raise RuntimeError("PlinkReader GRM calculation with subsetting not yet implemented. "
                               "Use PGEN/BGEN or full dataset.")
        _lib.check(_lib.lib().cb_grm_compute(self.handle, reader.handle, outputs))
        self._n = reader.getNumSamples()
        self._m = reader.getNumSnps()

    def getGrm(self) -> np.ndarray:
        G = np.empty((self._n, self._n), dtype=np.float64)
        _lib.check(_lib.lib().cb_grm_copy_grm(self.handle, _lib.ptr(G), self._n))
        return G

    def getGrmF32(self) -> np.ndarray:
        G = np.empty((self._n, self._n), dtype=np.float32)
        _lib.check(_lib.lib().cb_grm_copy_grm_f32(self.handle, _lib.ptr(G), self._n))
        return G

    def getIntGram(self) -> np.ndarray:
        G = np.empty((self._n, self._n), dtype=np.int32)
        _lib.check(_lib.lib().cb_grm_copy_int_gram(self.handle, _lib.ptr(G), self._n))
        return G

    def copyGrmToDevice(self, dptr: int, ld: int) -> None:
        """FP64 G into device memory owned by the caller (same device)."""
        _lib.check(_lib.lib().cb_grm_copy_grm_device(self.handle, dptr, ld))

    def devicePtr(self, which: int = _lib.CB_GRM_OUT_F64) -> tuple[int, int]:
        p = C.c_void_p()
        ld = C.c_int64()
        _lib.check(_lib.lib().cb_grm_device_ptr(self.handle, which, C.byref(p), C.byref(ld)))
        return p.value, ld.value

    def writeMatrix(self, path: str, fmt: str = "bin") -> None:
        """The FP64 G to `path`, streamed from the device, as
        write_matrix_txt ("txt") or write_upper_packed_binary ("bin")
        writes it (src/core/matrix_io.cpp:464-479)."""
        from .matrixio import FORMATS

        _lib.check(_lib.lib().cb_grm_write(self.handle, str(path).encode(), FORMATS[fmt]))

    def getSum2pq(self) -> float:
        v = C.c_double()
        _lib.check(_lib.lib().cb_grm_sum2pq(self.handle, C.byref(v)))
        return v.value

    def getTotalSnps(self) -> int:
        v = C.c_int64()
        _lib.check(_lib.lib().cb_grm_total_snps(self.handle, C.byref(v)))
        return v.value

    def getAlleleFreqs(self) -> np.ndarray:
        f = np.empty(self._m, dtype=np.float32)
        _lib.check(_lib.lib().cb_grm_copy_allele_freqs(self.handle, _lib.ptr(f), self._m))
        return f

    def getSnpStats(self) -> dict[str, np.ndarray]:
        m = self._m
        out = {k: np.empty(m, dtype=np.int32) for k in ("n_obs", "sum_x", "n_hom1", "n_het", "n_hom2")}
        out["valid"] = np.empty(m, dtype=np.uint8)
        out["mean_f"] = np.empty(m, dtype=np.float32)
        out["twopq_f"] = np.empty(m, dtype=np.float32)
        keys = ("n_obs", "sum_x", "n_hom1", "n_het", "n_hom2", "valid", "mean_f", "twopq_f")
        _lib.check(_lib.lib().cb_grm_copy_snp_stats(self.handle, m, *[_lib.ptr(out[k]) for k in keys]))
        return out
