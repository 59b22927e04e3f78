"""PlinkReader — host mirror of src/core/plink_reader.{h,cpp} for the BED path.

Reads .fam/.bim as whitespace tables (plink_reader.cpp:33-64), validates the
.bed magic (plink_reader.cpp:66-91) and hands the bytes to the device
(cb_genotypes_from_bed). Errors are RuntimeError with the reference's text.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _lib


@dataclass
class PlinkFamInfo:
    fid: str
    iid: str
    father: str
    mother: str
    sex: str
    phenotype: str


@dataclass
class PlinkBimInfo:
    chrom: str
    id: str
    cm_pos: str
    bp_pos: str
    alt: str
    ref: str


def _read_table(path: str, what: str, ncols: int) -> list[list[str]]:
    p = Path(path)
    if not p.exists():
        raise RuntimeError(f"Cannot open {what} file: {path}")
    rows = []
    with open(p) as f:
        for line in f:
            if not line.strip():
                continue
            toks = line.split()
            toks += [""] * (ncols - len(toks))
            rows.append(toks[:ncols])
    return rows


class PlinkReader:
    """load(prefix) reads prefix.{fam,bim,bed}; the genotypes live on the GPU."""

    def __init__(self, ctx: _lib.Context | None = None):
        self.ctx = ctx or _lib.default_context()
        self.fam_data: list[PlinkFamInfo] = []
        self.bim_data: list[PlinkBimInfo] = []
        self.handle = None
        self.bed_bytes: np.ndarray | None = None

    # plink_reader.cpp:27-31
    def load(self, prefix: str) -> None:
        self.readFam(prefix + ".fam")
        self.readBim(prefix + ".bim")
        self.openBed(prefix + ".bed")

    def readFam(self, path: str) -> None:
        self.fam_data = [PlinkFamInfo(*r) for r in _read_table(path, ".fam", 6)]
        if not self.fam_data:
            raise RuntimeError("No samples found in .fam file")

    def readBim(self, path: str) -> None:
        self.bim_data = [PlinkBimInfo(*r) for r in _read_table(path, ".bim", 6)]

    def openBed(self, path: str) -> None:
        if not Path(path).exists():
            raise RuntimeError(f"Cannot open .bed file: {path}")
        data = np.fromfile(path, dtype=np.uint8)
        self.from_bytes(data, len(self.fam_data), len(self.bim_data))

    def from_bytes(self, data: np.ndarray, n: int, m: int) -> None:
        """Take an in-memory .bed image (with magic) for n samples, m SNPs."""
        data = np.ascontiguousarray(data, dtype=np.uint8)
        h = C.c_void_p()
        _lib.check(_lib.lib().cb_genotypes_from_bed(self.ctx.handle, _lib.ptr(data), data.size, n, m,
                                                    C.byref(h)))
        self.close()
        self.handle = h
        self.bed_bytes = data
        self._n, self._m = n, m

    @classmethod
    def from_device_rows(cls, ctx: _lib.Context, d_ptr: int, row_pitch: int, n: int, m: int) -> "PlinkReader":
        r = cls(ctx)
        h = C.c_void_p()
        _lib.check(_lib.lib().cb_genotypes_from_device_rows(ctx.handle, d_ptr, row_pitch, n, m, C.byref(h)))
        r.handle = h
        r._n, r._m = n, m
        return r

    def close(self) -> None:
        if self.handle:
            _lib.lib().cb_genotypes_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def getNumSamples(self) -> int:
        return self._n

    def getNumSnps(self) -> int:
        return self._m

    def getBytesPerSnp(self) -> int:
        return (self._n + 3) // 4

    def getFamInfo(self) -> list[PlinkFamInfo]:
        return self.fam_data

    def getBimInfo(self) -> list[PlinkBimInfo]:
        return self.bim_data
