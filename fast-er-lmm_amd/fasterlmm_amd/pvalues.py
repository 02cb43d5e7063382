"""Caller-side statistics on the GPU (reference fasterlmm/cli.py:128-132, 405-412): -log10 p of the Wald F,
the permutation threshold and EffectSize."""
from __future__ import annotations

import numpy as np
import torch
from torch import Tensor

from . import _lib


def neglog10p(F: Tensor, df2: float) -> Tensor:
    """-log10 F(1, df2).sf(F) elementwise (cli.py:405-406), FP64, finite where p underflows FP64."""
    dev = _lib.device()
    home = F.device
    Fd = F.to(device=dev, dtype=torch.float64).contiguous()
    out = torch.empty_like(Fd)
    _lib.call("flmm_neglog10p_f1", Fd.data_ptr(), Fd.numel(), float(df2), out.data_ptr(), _lib.stream())
    return out.to(home)


def pvalue(F: Tensor, df2: float) -> Tensor:
    """p = 10 ** -(-log10 p) (underflows to 0 like scipy's sf for very large F)."""
    return torch.pow(10.0, -neglog10p(F, df2))


def perm_threshold_p(perm_max_F: Tensor, df2: float, quantile: float = 0.05) -> np.ndarray:
    """Per-phenotype genome-wide threshold: the linear-interpolated quantile of the per-permutation minimum p,
    p = F(1, df2).sf(max F) (cli.py:128, 406)."""
    p = pvalue(perm_max_F, df2).cpu().numpy()
    return np.quantile(p, quantile, axis=-1)


def effect_size(beta: Tensor, g_var: Tensor, p_var: Tensor) -> Tensor:
    """EffectSize[m, j] = beta^2 * g_var[m] / p_var[j] (cli.py:132) for an (M, n_real) beta."""
    dev = _lib.device()
    home = beta.device
    b = beta.to(device=dev, dtype=torch.float64).contiguous()
    g = torch.as_tensor(g_var).to(device=dev, dtype=torch.float64).contiguous()
    p = torch.as_tensor(p_var).to(device=dev, dtype=torch.float64).contiguous()
    M, n_real = b.shape
    out = torch.empty_like(b)
    _lib.call("flmm_effect_size", b.data_ptr(), M, n_real, g.data_ptr(), p.data_ptr(), out.data_ptr(), _lib.stream())
    return out.to(home)
