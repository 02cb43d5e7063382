"""Streamed / resident LOCO engine of the low-rank path (the reference's fasterlmm/extreme_scan.py) on MI355X.

One shared chromosome loop (_loco_scan_core, extreme_scan.py:28-77): per chromosome the kinship factor drops that
chromosome's markers and is decomposed once (optionally truncated to a rank with the randomised range finder),
the columns are rotated and their deltas fitted once, the column-side scan planes are built once, and then every
test block of the chromosome streams past them.  The two entry points differ only in where the blocks come from:
``loco_scan_streamed`` decodes them from a memory-mapped .bed on the GPU block by block (flmm_bed_decode with
fused standardisation), ``loco_scan_resident`` slices a genotype matrix already in HBM (FP64 standardised, or raw
int8 dosages standardised per block on the GPU -- per-column standardisation makes the two identical).
"""
from __future__ import annotations

import numpy as np
import torch
from torch import Tensor

from . import engine
from .core import ScanResult, _dev, _home
from .lowrank import (LowRankSpectrum, SplitFactor, _basis, _ColumnPlanes, _scan_blocks, _state,
                      fit_delta_grid_lowrank, refresh_real_max, rotate_phenos)


def _loco_scan_core(chroms: list, block_source, M: int, G: Tensor, g_chrom: np.ndarray, X: Tensor, Y: Tensor, *,
                    n_real: int | None, block_size: int, on_chrom=None, rank: int | None = None,
                    dtype: torch.dtype | None = None) -> ScanResult:
    """extreme_scan.py:28-77.  block_source(c) yields (standardised (N, b) block, global column indices)."""
    _, (Gd, Xd, Yd) = _dev(G, X, Y)
    g_chrom = np.asarray(g_chrom)
    P = Yd.shape[1]
    n_real = P if n_real is None else min(int(n_real), P)
    dev = Gd.device
    out = engine.ScanOutputs.allocate(M, n_real, P, dev)
    # the range finder's products run against G's MFMA operand planes, packed once for every chromosome
    factor = SplitFactor(Gd, engine.prec_code()) if rank is not None else None
    for k_idx, c in enumerate(chroms):
        with engine.stage("basis"):
            keep = torch.from_numpy((g_chrom != c).astype(np.float64)).to(dev)
            basis = _basis(Gd, Xd, keep, rank=rank, factor=factor)
            del keep
        with engine.stage("rotate_columns"):
            UY, UUY = rotate_phenos(basis, Yd)
            spec = LowRankSpectrum(s=basis.s, U_eff=basis.U_eff, X=basis.X, Xpinv=basis.Xpinv, Neff=basis.Neff,
                                   UY=UY, UUY=UUY)
        with engine.stage("delta"):
            ld = fit_delta_grid_lowrank(spec)
            st = _state(spec)
            cp = _ColumnPlanes(st, spec, st.delta_for(ld))
        for S_b, cidx in block_source(c):
            S_b = S_b.to(device=dev, dtype=torch.float64)
            ridx = torch.from_numpy(np.asarray(cidx, dtype=np.int64)).to(dev)
            _scan_blocks(st, spec, cp, S_b, ridx, n_real, out, block_size)
        del spec, st, cp, UY, UUY, basis
        if on_chrom is not None:
            on_chrom(k_idx + 1, len(chroms))
    refresh_real_max(out, n_real)
    return ScanResult(out.f, out.beta, out.se, out.sfve, out.nullh2, out.max_F)


def loco_scan_streamed(handle, G: Tensor, g_chrom: np.ndarray, X: Tensor, Y: Tensor, *, n_real: int | None = None,
                       block_size: int = 8192, dtype: torch.dtype = torch.float32, on_chrom=None,
                       rank: int | None = None) -> ScanResult:
    """extreme_scan.py:83-103: test variants decoded off an aligned .bed (bed.open_aligned_bed) block by block on
    the GPU against a resident kinship factor G (N, k) with per-column chromosome labels g_chrom.  The blocks are
    decoded in FP64 whatever `dtype` says (it is accepted for signature compatibility)."""
    from .bed import chrom_test_blocks
    home = Y.device if isinstance(Y, Tensor) else torch.device("cpu")
    chroms = sorted(set(np.asarray(handle.chrom).tolist()))

    def block_source(c):
        return chrom_test_blocks(handle, c, block_size=block_size, dtype=torch.float64)

    res = _loco_scan_core(chroms, block_source, len(handle.sid), G, g_chrom, X, Y, n_real=n_real,
                          block_size=block_size, on_chrom=on_chrom, rank=rank)
    return _home(res, home)


def loco_scan_resident(Z: Tensor, z_chrom: np.ndarray, G: Tensor, g_chrom: np.ndarray, X: Tensor, Y: Tensor, *,
                       n_real: int | None = None, block_size: int = 8192, dtype: torch.dtype = torch.float32,
                       on_chrom=None, rank: int | None = None) -> ScanResult:
    """extreme_scan.py:106-125: test blocks sliced from a resident genotype matrix Z (N, M): FP64/FP32 already
    standardised, or int8 raw dosages (negative = missing) standardised block by block on the GPU."""
    from .io import standardise_columns
    home = Y.device if isinstance(Y, Tensor) else torch.device("cpu")
    z_chrom = np.asarray(z_chrom)
    Zt = torch.as_tensor(Z)
    raw = Zt.dtype == torch.int8
    chroms = sorted(set(z_chrom.tolist()))

    def block_source(c):
        cols = np.flatnonzero(z_chrom == c)
        for s in range(0, len(cols), block_size):
            cidx = cols[s:s + block_size]
            blk = Zt[:, torch.from_numpy(cidx).to(Zt.device)]
            yield (standardise_columns(blk) if raw else blk), cidx

    res = _loco_scan_core(chroms, block_source, Zt.shape[1], G, g_chrom, X, Y, n_real=n_real,
                          block_size=block_size, on_chrom=on_chrom, rank=rank)
    return _home(res, home)
