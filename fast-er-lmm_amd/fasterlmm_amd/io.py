"""Drop-in scan-input helpers of the reference's fasterlmm/io.py on MI355X: the aligned dataset, column
standardisation (a gfx950 kernel), the kinship Gram, NA-pattern grouping and strain subsets.

File readers (PLINK / phenotype / covariate tables) are not part of this module; ``AlignedDataset`` is built
from arrays with ``align_inputs``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch
from torch import Tensor

from . import _lib


@dataclass
class AlignedDataset:
    """Strain-aligned bundle (io.py:123-140): Z (N, M) raw genotypes, Y (N, P) phenotypes, X (N, C) =
    [intercept | covariates], chrom/pos/snp_id per SNP, optional separate kinship panel Z_kin/chrom_kin."""
    iid: list
    Z: Tensor
    Y: Tensor
    X: Tensor
    chrom: np.ndarray
    pos: np.ndarray
    snp_id: list
    pheno_names: list
    Z_kin: Tensor | None = None
    chrom_kin: np.ndarray | None = None
    n_dropped_kin: int = 0


def align_inputs(geno_iid, Z, pheno_iid, Y, covar_iid=None, C=None, *, chrom, pos=None, snp_id=None,
                 pheno_names=None, dtype=torch.float64, kin_iid=None, Z_kin=None, chrom_kin=None) -> AlignedDataset:
    """Strain intersection in genotype order and X = [1 | covariates] (io.py:143-184), from arrays."""
    common = set(geno_iid) & set(pheno_iid)
    if covar_iid is not None:
        common &= set(covar_iid)
    n_dropped_kin = 0
    if kin_iid is not None:
        n_before = len(common)
        common &= set(kin_iid)
        n_dropped_kin = n_before - len(common)
    if not common:
        raise ValueError("no strain overlap between geno / pheno / covar / kinship")
    iid = [s for s in geno_iid if s in common]
    gi = {s: i for i, s in enumerate(geno_iid)}
    pi = {s: i for i, s in enumerate(pheno_iid)}
    Zn = np.asarray(Z)[[gi[s] for s in iid], :]
    Yn = np.asarray(Y, dtype=np.float64)[[pi[s] for s in iid], :]
    if covar_iid is not None:
        ci = {s: i for i, s in enumerate(covar_iid)}
        X = np.concatenate([np.ones((len(iid), 1)), np.asarray(C, dtype=np.float64)[[ci[s] for s in iid], :]], axis=1)
    else:
        X = np.ones((len(iid), 1))
    Zk = None
    if kin_iid is not None:
        ki = {s: i for i, s in enumerate(kin_iid)}
        Zk = torch.from_numpy(np.asarray(Z_kin)[[ki[s] for s in iid], :]).to(dtype)
    M = Zn.shape[1]
    return AlignedDataset(iid=iid, Z=torch.from_numpy(Zn).to(dtype), Y=torch.from_numpy(Yn).to(dtype),
                          X=torch.from_numpy(X).to(dtype), chrom=np.asarray(chrom),
                          pos=np.arange(M) if pos is None else np.asarray(pos),
                          snp_id=[str(i) for i in range(M)] if snp_id is None else list(snp_id),
                          pheno_names=[f"p{i}" for i in range(Yn.shape[1])] if pheno_names is None else list(pheno_names),
                          Z_kin=Zk, chrom_kin=None if chrom_kin is None else np.asarray(chrom_kin),
                          n_dropped_kin=n_dropped_kin)


def group_by_na(Y: Tensor, cols: list[int]) -> list[tuple[list[int], list[int]]]:
    """(keep_rows, cols) per distinct NA pattern, first-seen order (io.py:187-203)."""
    mask = torch.isnan(torch.as_tensor(Y)[:, cols]).cpu().numpy()
    groups: dict[bytes, list[int]] = {}
    for j, _ in enumerate(cols):
        groups.setdefault(np.packbits(mask[:, j]).tobytes(), []).append(j)
    out = []
    for js in groups.values():
        keep = np.flatnonzero(~mask[:, js[0]]).tolist()
        out.append((keep, [cols[j] for j in js]))
    return out


def subset_dataset(data: AlignedDataset, rows: list[int], cols: list[int]) -> AlignedDataset:
    """Strain and phenotype subset; genotypes keep their raw missing calls (io.py:206-218)."""
    r = torch.as_tensor(rows, dtype=torch.long, device=data.Z.device)
    return AlignedDataset(iid=[data.iid[i] for i in rows], Z=data.Z[r], Y=data.Y[r][:, cols], X=data.X[r],
                          chrom=data.chrom, pos=data.pos, snp_id=data.snp_id,
                          pheno_names=[data.pheno_names[c] for c in cols],
                          Z_kin=data.Z_kin[r] if data.Z_kin is not None else None,
                          chrom_kin=data.chrom_kin, n_dropped_kin=data.n_dropped_kin)


def standardise_columns(arr, *, stats: bool = False):
    """Unit standardisation per column on the GPU (io.py:223-254): NaN-aware mean and population std, (x - mean)
    / std, std == 0 -> the column becomes 0, NaN -> 0 after.  Accepts FP64/FP32 (NaN = missing) or int8 (< 0 =
    missing) tensors or numpy arrays; returns the same container type (FP64) on the input's device.  A 1-D input
    is one column.  stats=True also returns (mean, std)."""
    is_np = isinstance(arr, np.ndarray)
    t = torch.from_numpy(np.ascontiguousarray(arr)) if is_np else arr
    home = t.device
    dev = _lib.device()
    one_d = t.ndim == 1
    t2 = t.reshape(-1, 1) if one_d else t
    if t2.dtype == torch.int8:
        src, code = t2.to(dev).contiguous(), 1
    else:
        src, code = t2.to(device=dev, dtype=torch.float64).contiguous(), 0
    n, m = src.shape
    dst = torch.empty(n, m, dtype=torch.float64, device=dev)
    mean = torch.empty(m, dtype=torch.float64, device=dev)
    std = torch.empty(m, dtype=torch.float64, device=dev)
    _lib.call("flmm_standardise", code, src.data_ptr(), n, m, m, dst.data_ptr(), m, mean.data_ptr(), std.data_ptr(),
              _lib.stream())
    out = dst.reshape(-1) if one_d else dst
    out = out.cpu().numpy() if is_np else out.to(home)
    if stats:
        return out, (mean, std)
    return out


def grm(Z: Tensor) -> Tensor:
    """K = Z Z^T / M on pre-standardised Z (io.py:257-265); FP64 on the GPU."""
    return (Z @ Z.T) / Z.shape[1]
