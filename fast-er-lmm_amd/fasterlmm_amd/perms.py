"""Drop-in permutation driver (reference fasterlmm/perms.py:21-68) on MI355X.

The permutation orders come from numpy's PCG64 generator seeded per (seed, phenotype column), exactly as the
reference draws them (perms.py:51-55), so a phenotype sees the same shuffles whatever batch it rides in.  The
packed [reals | b0 perms | b1 perms | ...] matrix is gathered on the GPU and scanned in one compat LOCO (or
single-kinship) pass.
"""
from __future__ import annotations

import numpy as np
import torch
from torch import Tensor

from . import _lib
from .core import ScanResult, loco_scan_compat, single_k_scan_compat
from .io import AlignedDataset, standardise_columns


def perm_orders(seed: int, seed_cols, N: int, n_perm: int) -> np.ndarray:
    """orders[n, b, j] = row landing in position n of column b's j-th shuffle (perms.py:51-55)."""
    seed_cols = list(seed_cols)
    orders = np.empty((N, len(seed_cols), n_perm), dtype=np.int64)
    for b, p in enumerate(seed_cols):
        rng = np.random.default_rng([seed, int(p)])
        orders[:, b, :] = rng.random((n_perm, N)).argsort(axis=1).T
    return orders


def pack_perm_columns(y_real: Tensor, orders: np.ndarray) -> Tensor:
    """(N, B + B*n_perm) pheno-major packing of the reals and their shuffles (perms.py:56-59)."""
    N, B = y_real.shape
    n_perm = orders.shape[2]
    gi = torch.from_numpy(orders).to(y_real.device)
    y_perms = torch.gather(y_real.unsqueeze(2).expand(N, B, n_perm), 0, gi).reshape(N, B * n_perm)
    return torch.cat([y_real, y_perms], dim=1)


def perm_threshold(data: AlignedDataset, pheno_idx: list[int], *, n_perm: int = 100, seed: int = 19930909,
                   loco: bool = True, on_chrom=None, seed_idx: list[int] | None = None) -> tuple[ScanResult, Tensor]:
    """Scan B real phenotypes and their B * n_perm shuffles together; returns (ScanResult of the B reals,
    perm_max_F (B, n_perm)) (perms.py:21-68)."""
    dev = _lib.device()
    home = data.Y.device
    Z_std = standardise_columns(data.Z.to(dev))
    Z_kin_std = standardise_columns(data.Z_kin.to(dev)) if data.Z_kin is not None else None
    B = len(pheno_idx)
    y_real = data.Y[:, pheno_idx].to(device=dev, dtype=torch.float64)
    N = y_real.shape[0]
    orders = perm_orders(seed, pheno_idx if seed_idx is None else seed_idx, N, n_perm)
    Y_all = pack_perm_columns(y_real, orders)
    X = data.X.to(device=dev, dtype=torch.float64)
    if loco:
        res = loco_scan_compat(Z_std, X, Y_all, data.chrom, n_real=B, on_chrom=on_chrom, Z_kin=Z_kin_std,
                               chrom_kin=data.chrom_kin)
    else:
        res = single_k_scan_compat(Z_std, X, Y_all, n_real=B, Z_kin=Z_kin_std)
    perm_max_F = res.max_F[B:].reshape(B, n_perm)
    if home.type != "cuda":
        res = ScanResult(*(getattr(res, k).to(home) for k in ("f", "beta", "se", "sfve", "nullh2", "max_F")))
        perm_max_F = perm_max_F.to(home)
    return res, perm_max_F


def perm_threshold_groups(data: AlignedDataset, pheno_idx: list[int], *, n_perm: int = 100, seed: int = 19930909,
                          mode: str = "grouped", runner=None, on_chrom=None, test_chroms=None,
                          loco: bool = True) -> list:
    """The reference's NA-grouped permutation run (cli.py:337-375 with perms.py:21-68 per group) with the scans of
    every group fused: phenotypes are grouped by NA pattern (io.group_by_na), each group keeps its observed
    strains and is scanned with its reals and their shuffles (seeded by the full-table column, as the reference's
    ``seed_idx``), all groups through one grouped launch per SNP block (mode="grouped") or one launch pair per
    group (mode="loop").  Groups with fewer than C + 2 strains are skipped, as in the reference.

    Returns [(keep_rows, cols, ScanResult of the group's reals, perm_max_F (B_g, n_perm))] per scanned group, equal
    to ``perm_threshold(subset_dataset(data, keep, cols), range(len(cols)), seed_idx=cols, loco=loco)`` per group
    (loco=False: one kinship of all SNPs per group)."""
    from .grouped import loco_scan_compat_groups
    from .io import group_by_na
    dev = _lib.device()
    C = data.X.shape[1]
    groups = [(keep, cols) for keep, cols in group_by_na(data.Y, pheno_idx) if len(keep) - C - 1 >= 1]
    if not groups:
        return []
    Ys, keeps, n_reals = [], [], []
    for keep, cols in groups:
        r = torch.as_tensor(keep, dtype=torch.long, device=data.Y.device)
        y_real = data.Y[r][:, cols].to(device=dev, dtype=torch.float64)
        orders = perm_orders(seed, cols, len(keep), n_perm)
        Ys.append(pack_perm_columns(y_real, orders))
        keeps.append(keep)
        n_reals.append(len(cols))
    res = loco_scan_compat_groups(data.Z, data.X, Ys, keeps, data.chrom, n_reals=n_reals, Z_kin=data.Z_kin,
                                  chrom_kin=data.chrom_kin, mode=mode, runner=runner, on_chrom=on_chrom,
                                  test_chroms=test_chroms, loco=loco)
    out = []
    home = data.Y.device
    for (keep, cols), r in zip(groups, res):
        B = len(cols)
        pmf = r.max_F[B:].reshape(B, n_perm)
        if home.type != "cuda":
            r = ScanResult(*(getattr(r, k).to(home) for k in ("f", "beta", "se", "sfve", "nullh2", "max_F")))
            pmf = pmf.to(home)
        out.append((keep, cols, r, pmf))
    return out
