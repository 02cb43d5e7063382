"""GPU -log10 p / EffectSize against scipy and mpmath (the caller math of cli.py:128-132, 405-412)."""
from __future__ import annotations

import mpmath
import numpy as np
import pytest
import torch
from scipy import stats

pytestmark = pytest.mark.gpu


def test_neglog10p_matches_scipy_and_mpmath():
    from fasterlmm_amd.pvalues import neglog10p
    rng = np.random.default_rng(0)
    for df in (3.0, 146.0, 995.0, 4993.0):
        F = np.concatenate([[0.0, 1e-12, 1e-3, 0.5, 1.0, 2.0, 3.84, 10.0, 50.0, 300.0, 2000.0],
                            rng.gamma(1.0, 5.0, size=2000)])
        got = neglog10p(torch.from_numpy(F).cuda(), df).cpu().numpy()
        want = -stats.f.logsf(F, 1, df) / np.log(10)
        ok = np.isfinite(want)
        assert np.abs(got[ok] - want[ok]).max() <= 1e-9 * np.maximum(1, np.abs(want[ok])).max()
    # deep tail, beyond the FP64 range of p itself: against mpmath's regularised incomplete beta
    for F, df in ((5000.0, 995.0), (20000.0, 995.0), (1e5, 146.0)):
        got = neglog10p(torch.tensor([F], dtype=torch.float64).cuda(), df).item()
        exact = -float(mpmath.log10(mpmath.betainc(df / 2.0, 0.5, 0, df / (df + F), regularized=True)))
        assert abs(got - exact) <= 1e-9 * exact
    edge = neglog10p(torch.tensor([float("nan"), float("inf"), -1.0], dtype=torch.float64).cuda(), 10.0).cpu()
    assert np.isnan(edge[0]) and np.isinf(edge[1]) and edge[2] == 0.0


def test_effect_size_and_threshold():
    from fasterlmm_amd.pvalues import effect_size, perm_threshold_p
    rng = np.random.default_rng(1)
    beta = rng.normal(size=(50, 3))
    g = rng.uniform(0.1, 0.5, size=50)
    p = rng.uniform(0.5, 2.0, size=3)
    got = effect_size(torch.from_numpy(beta).cuda(), torch.from_numpy(g), torch.from_numpy(p)).cpu().numpy()
    np.testing.assert_allclose(got, beta ** 2 * g[:, None] / p[None, :], rtol=1e-15)
    pmf = rng.gamma(3.0, 5.0, size=(3, 100))
    thr = perm_threshold_p(torch.from_numpy(pmf).cuda(), 146.0)
    np.testing.assert_allclose(thr, np.quantile(stats.f.sf(pmf, 1, 146.0), 0.05, axis=-1), rtol=1e-9)
