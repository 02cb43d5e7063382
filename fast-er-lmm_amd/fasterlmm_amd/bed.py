"""PLINK 1 .bed/.bim/.fam reading with the decode and the unit standardisation on the GPU.

Keeps the reference's reader API -- ``Genotypes`` / ``read_plink`` (fasterlmm/io.py:24-35, 54-67) and the
streaming helpers of fasterlmm/io_stream.py (``open_aligned_bed``, ``read_block``, ``select_grm_markers``,
``read_grm_factor``, ``stream_genotype_var``, ``chrom_test_blocks``, ``read_all_standardised``,
``resident_chrom_blocks``: io_stream.py:40-170) -- without pysnptools: the .bim/.fam text is parsed on the host,
the packed 2-bit records are copied to the device as they are, and ``flmm_bed_decode`` expands (and, for the
blocks the scan consumes, standardises) them there.  Blocks come back on the GPU in the requested dtype.
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch
from torch import Tensor

from . import _lib

MAGIC = bytes([0x6C, 0x1B, 0x01])


@dataclass
class Genotypes:
    """Plink BED contents (io.py:24-35): Z (N, M) count_A1 dosages 0/1/2 as float64 with NaN for missing, iid =
    N strain IDs, sid = M variant IDs, chrom and pos per variant."""
    Z: np.ndarray
    iid: list[str]
    sid: list[str]
    chrom: np.ndarray
    pos: np.ndarray


# PLINK's non-numeric chromosome codes (human numbering)
_CHROM_CODES = {"X": 23.0, "Y": 24.0, "XY": 25.0, "MT": 26.0, "M": 26.0}


def _chrom_value(label: str) -> float:
    lab = label.upper().removeprefix("CHR")
    return _CHROM_CODES[lab] if lab in _CHROM_CODES else float(lab)


def _read_text(prefix: Path):
    """.fam -> strain IDs (column 2); .bim -> variant IDs (column 2), chromosome (column 1) and base-pair
    position (column 4), as pysnptools exposes them (Bed.iid[:, 1], Bed.sid, Bed.pos[:, 0], Bed.pos[:, 2])."""
    fam = [ln.split() for ln in prefix.with_suffix(".fam").read_text().splitlines() if ln.strip()]
    bim = [ln.split() for ln in prefix.with_suffix(".bim").read_text().splitlines() if ln.strip()]
    iid = [r[1] for r in fam]
    sid = [r[1] for r in bim]
    chrom = np.array([_chrom_value(r[0]) for r in bim])
    pos = np.array([float(r[3]) for r in bim])
    return iid, sid, chrom, pos


def _open_bed(prefix: Path, n: int, m: int) -> np.ndarray:
    raw = np.memmap(prefix.with_suffix(".bed"), dtype=np.uint8, mode="r")
    if raw.size < 3 or bytes(raw[:3]) != MAGIC:
        raise ValueError(f"{prefix.with_suffix('.bed')}: not a SNP-major PLINK 1 .bed")
    rec = (n + 3) // 4
    if raw.size < 3 + rec * m:
        raise ValueError(f"{prefix.with_suffix('.bed')}: {raw.size} bytes, expected {3 + rec * m}")
    return raw[3:3 + rec * m].reshape(m, rec)


def _decode(records: np.ndarray, n_file: int, row_idx, kind: int, count_A1: bool, dev, *, stats=False):
    """Device decode of the (b, rec) host records -> (n, b) tensor of the given kind (+ mean, std)."""
    b, rec = records.shape
    recs = torch.from_numpy(np.ascontiguousarray(records)).to(dev)
    if row_idx is None:
        n, rows = n_file, None
    else:
        rows = torch.as_tensor(np.asarray(row_idx, dtype=np.int64)).to(dev)
        n = rows.numel()
    dt = {_lib.BED_I8: torch.int8, _lib.BED_F64: torch.float64, _lib.BED_F32: torch.float32}[kind]
    out = torch.empty((n, b), dtype=dt, device=dev)
    mean = torch.empty(b, dtype=torch.float64, device=dev) if stats else None
    std = torch.empty(b, dtype=torch.float64, device=dev) if stats else None
    _lib.call("flmm_bed_decode", recs.data_ptr(), rec, rows.data_ptr() if rows is not None else None, n, None, b,
              int(bool(count_A1)), kind, out.data_ptr(), b, mean.data_ptr() if stats else None,
              std.data_ptr() if stats else None, _lib.stream())
    return (out, mean, std) if stats else out


def read_plink(prefix, *, count_A1: bool = True) -> Genotypes:
    """Read a plink BED/BIM/FAM trio (io.py:54-67).  Missing dosages stay NaN."""
    prefix = Path(prefix)
    iid, sid, chrom, pos = _read_text(prefix)
    recs = _open_bed(prefix, len(iid), len(sid))
    g = _decode(recs, len(iid), None, _lib.BED_I8, count_A1, _lib.device())
    Z = g.cpu().numpy().astype(np.float64)
    Z[Z < 0] = np.nan
    return Genotypes(Z=Z, iid=iid, sid=sid, chrom=chrom, pos=pos)


@dataclass
class BedHandle:
    """An opened .bed aligned to a strain set (io_stream.py:23-37): records stay on disk (memory-mapped)."""
    records: np.ndarray  # (M, ceil(N_file / 4)) uint8, memory-mapped
    n_file: int
    row_idx: np.ndarray
    iid: list[str]
    sid: list[str]
    chrom: np.ndarray
    pos: np.ndarray
    count_A1: bool = True


def open_aligned_bed(prefix, keep_iid: list[str], *, count_A1: bool = True) -> BedHandle:
    """Open a .bed lazily with its rows aligned to keep_iid, in the bed's own row order filtered to keep_iid
    (io_stream.py:40-56)."""
    prefix = Path(prefix)
    bed_iid, sid, chrom, pos = _read_text(prefix)
    common = set(keep_iid)
    row_idx = np.array([i for i, s in enumerate(bed_iid) if s in common], dtype=np.int64)
    return BedHandle(records=_open_bed(prefix, len(bed_iid), len(sid)), n_file=len(bed_iid), row_idx=row_idx,
                     iid=[bed_iid[i] for i in row_idx], sid=sid, chrom=chrom, pos=pos, count_A1=count_A1)


def read_block(h: BedHandle, col_idx, *, dtype: torch.dtype = torch.float32, device=None) -> Tensor:
    """One block of variants for the aligned strains, standardised (io_stream.py:59-71): an (N, len(col_idx))
    tensor on the GPU.  The standardisation is per column, so a block is identical whether standardised alone or
    with the whole matrix.  float64 and float32 are produced directly by the decoder; other dtypes are cast from
    float64."""
    dev = _lib.device() if device is None else torch.device(device)
    recs = h.records[np.asarray(col_idx, dtype=np.int64)]
    kind = _lib.BED_F32 if dtype == torch.float32 else _lib.BED_F64
    out = _decode(recs, h.n_file, h.row_idx, kind, h.count_A1, dev)
    return out if out.dtype == dtype else out.to(dtype)


def select_grm_markers(h: BedHandle, k: int, *, seed: int = 19930909) -> np.ndarray:
    """About k kinship markers by an even stride across the genome (io_stream.py:77-88)."""
    M = len(h.sid)
    if k >= M:
        return np.arange(M, dtype=np.int64)
    return np.unique(np.linspace(0, M - 1, k).round().astype(np.int64))


def read_grm_factor(h: BedHandle, col_idx, *, dtype: torch.dtype = torch.float32) -> Tensor:
    """The pruned kinship markers as a resident standardised (N, k) factor (io_stream.py:91-98)."""
    return read_block(h, col_idx, dtype=dtype)


def stream_genotype_var(h: BedHandle, *, block_size: int = 8192) -> np.ndarray:
    """Per-variant variance (NaN-aware, ddof 0) of the raw dosages over the aligned strains (io_stream.py:103-117),
    from the decoder's column statistics."""
    M = len(h.sid)
    out = np.empty(M, dtype=np.float64)
    dev = _lib.device()
    for s in range(0, M, block_size):
        cidx = np.arange(s, min(s + block_size, M))
        _, _, std = _decode(h.records[cidx], h.n_file, h.row_idx, _lib.BED_F64, h.count_A1, dev, stats=True)
        sd = std.cpu().numpy()
        out[cidx] = np.where(np.isinf(sd), 0.0, sd * sd)
    return out


def chrom_test_blocks(h: BedHandle, chrom_value, *, block_size: int = 8192, dtype: torch.dtype = torch.float32):
    """Yield (standardised (N, b) block, global column indices) for one chromosome (io_stream.py:120-132)."""
    cols = np.where(h.chrom == chrom_value)[0]
    for s in range(0, len(cols), block_size):
        cidx = cols[s:s + block_size]
        yield read_block(h, cidx, dtype=dtype), cidx


def read_all_standardised(h: BedHandle, *, block_size: int = 8192, dtype: torch.dtype = torch.float32) -> Tensor:
    """Decode + standardise every variant once into a resident (N, M) GPU tensor (io_stream.py:135-154)."""
    M, N = len(h.sid), len(h.iid)
    Z = torch.empty((N, M), dtype=dtype, device=_lib.device())
    for s in range(0, M, block_size):
        e = min(s + block_size, M)
        Z[:, s:e] = read_block(h, np.arange(s, e), dtype=dtype)
    return Z


def resident_chrom_blocks(Z: Tensor, chrom_labels: np.ndarray, chrom_value, *, block_size: int = 8192):
    """One chromosome's standardised blocks sliced from a resident (N, M) tensor (io_stream.py:157-170)."""
    cols = np.where(chrom_labels == chrom_value)[0]
    for s in range(0, len(cols), block_size):
        cidx = cols[s:s + block_size]
        yield Z.index_select(1, torch.as_tensor(cidx, device=Z.device)), cidx
