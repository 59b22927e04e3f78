"""ssGBLUP H^-1 — host mirror of the split-mode set-up of
src/io/relationship_builder.cpp:481-633 over the C-ABI (cb_hinv_blocks,
cb_mme_set_qinv_split). Device buffers are torch tensors (plumbing only)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def hinv_options(tune_g: bool = True, alpha: float = 0.05, tau: float = 1.0, omega: float = 1.0,
                 ridge: float = 0.0) -> _lib.HinvOptions:
    o = _lib.HinvOptions()
    _lib.lib().cb_hinv_options_default(C.byref(o))
    o.tune_g, o.alpha, o.tau, o.omega, o.ridge = int(tune_g), alpha, tau, omega, ridge
    return o


def build_hinv_blocks(dam, sire, gmap, G, opts: _lib.HinvOptions | None = None, want_a22: bool = False,
                      ctx: _lib.Context | None = None):
    """D = tau G^-1 - omega A22^-1 on the device for the genotyped rows
    `gmap` (pedigree row of each genotyped animal, -1 if absent). G is the
    ng x ng GRM (numpy or a CUDA float64 tensor). Returns (D, A22 or None,
    stats dict) with D and A22 as CUDA float64 tensors."""
    import torch

    ctx = ctx or _lib.default_context()
    opts = opts or hinv_options()
    dam = np.ascontiguousarray(dam, dtype=np.int32)
    sire = np.ascontiguousarray(sire, dtype=np.int32)
    gmap = np.ascontiguousarray(gmap, dtype=np.int32)
    ng = gmap.size
    if not torch.is_tensor(G):
        G = torch.from_numpy(np.ascontiguousarray(G, dtype=np.float64)).cuda()
    G = G.contiguous()
    D = torch.empty((ng, ng), dtype=torch.float64, device=G.device)
    A22 = torch.empty((ng, ng), dtype=torch.float64, device=G.device) if want_a22 else None
    stats = np.zeros(6)
    torch.cuda.synchronize()
    _lib.check(_lib.lib().cb_hinv_blocks(ctx.handle, dam.size, _lib.ptr(dam), _lib.ptr(sire), ng, _lib.ptr(gmap),
                                         G.data_ptr(), ng, C.byref(opts), D.data_ptr(), ng,
                                         A22.data_ptr() if want_a22 else None, ng, _lib.ptr(stats)))
    keys = ("a", "b", "mean_diag_A22", "mean_offdiag_A22", "mean_diag_G", "mean_offdiag_G")
    return D, A22, dict(zip(keys, stats.tolist()))
