"""FP64 CPU oracle for the permutation driver and the caller's F -> p math -- TEST INFRASTRUCTURE ONLY.

Restates fasterlmm/perms.py:21-68 and the caller arithmetic in fasterlmm/cli.py:128-132, 405-412.  Only
tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it.

Pinning: the permutation orders come from numpy's PCG64 ``default_rng([seed, p])`` (a library routine), and
the packing is checked against an independent per-column re-gather, the scheme the reference's own
tests/test_cli_extreme.py:48-66 pins.  F -> p is scipy's F survival function (cli.py:405-406), checked
against mpmath and the Student-t identity F(1, v) = t(v)^2.
"""
from __future__ import annotations

import numpy as np
from scipy import stats

from oracle import lmm


def perm_orders(seed: int, seed_cols, N: int, n_perm: int) -> np.ndarray:
    """orders[n, b, j] = row landing in position n of column b's j-th shuffle (perms.py:51-55)."""
    seed_cols = list(seed_cols)
    orders = np.empty((N, len(seed_cols), n_perm), dtype=np.int64)
    for b, p in enumerate(seed_cols):
        rng = np.random.default_rng([seed, int(p)])
        orders[:, b, :] = rng.random((n_perm, N)).argsort(axis=1).T
    return orders


def pack_perm_columns(y_real: np.ndarray, orders: np.ndarray) -> np.ndarray:
    """[reals | b0 j0..jP | b1 ...] pheno-major packing (perms.py:56-59)."""
    N, B = y_real.shape
    n_perm = orders.shape[2]
    y_perms = np.take_along_axis(np.broadcast_to(y_real[:, :, None], (N, B, n_perm)), orders, axis=0)
    return np.concatenate([y_real, y_perms.reshape(N, B * n_perm)], axis=1)


def perm_threshold(Z_raw, X, Y, chrom, pheno_idx, *, n_perm=100, seed=19930909, loco=True, seed_idx=None,
                   Z_kin_raw=None, chrom_kin=None, n_grid=64):
    """perm_threshold (perms.py:21-68) on plain arrays: standardise, pack reals + perms, compat scan,
    perm_max_F = max_F[B:].reshape(B, n_perm)."""
    Z = lmm.standardise_columns(Z_raw)
    Z_kin = lmm.standardise_columns(Z_kin_raw) if Z_kin_raw is not None else None
    B = len(pheno_idx)
    y_real = Y[:, pheno_idx]
    orders = perm_orders(seed, pheno_idx if seed_idx is None else seed_idx, y_real.shape[0], n_perm)
    Y_all = pack_perm_columns(y_real, orders)
    if loco:
        res = lmm.loco_scan_compat(Z, X, Y_all, chrom, n_real=B, Z_kin=Z_kin, chrom_kin=chrom_kin,
                                   n_grid=n_grid)
    else:
        res = lmm.single_k_scan_compat(Z, X, Y_all, n_real=B, Z_kin=Z_kin, n_grid=n_grid)
    return res, res["max_F"][B:].reshape(B, n_perm)


def f_pvalue(F, df2):
    """p = F(1, df2) survival function (cli.py:405-406)."""
    return stats.f.sf(F, 1, df2)


def neglog10_pvalue(F, df2):
    """-log10 p through the log survival function, finite where sf underflows."""
    return -stats.f.logsf(F, 1, df2) / np.log(10.0)


def perm_threshold_value(perm_max_F, df2, quantile=0.05):
    """Genome-wide threshold = linear quantile of the per-permutation minimum p (cli.py:128, 406)."""
    return np.quantile(f_pvalue(perm_max_F, df2), quantile, axis=-1)


def effect_size(beta, g_var, p_var):
    """EffectSize = beta^2 * var(genotype) / var(phenotype) (cli.py:132)."""
    return beta * beta * g_var / p_var
