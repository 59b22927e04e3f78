"""cosmicblup_amd — MI355X-native GBLUP/ssGBLUP hot path of CosmicBLUP.

Host mirrors of the reference interfaces over the C-ABI library
libcosmicblup_amd.so (include/cosmicblup_amd.h).
"""
from ._lib import CB_GRM_OUT_F32, CB_GRM_OUT_F64, CB_GRM_OUT_INT, Context, CosmicError, NotConverged
from .grm import GrmBuilder, Options
from .plink import PlinkReader

__all__ = [
    "Context", "CosmicError", "NotConverged", "GrmBuilder", "Options", "PlinkReader",
    "CB_GRM_OUT_F64", "CB_GRM_OUT_F32", "CB_GRM_OUT_INT",
]
