"""The factored weights' exponential sum 1 / z ~= sum_r lam_r^2 exp(-t_r z) (fasterlmm_amd/expsum.py), checked
against 1 / z itself on dense grids, and the factored sWs against the exact sum on small random spectra."""
from __future__ import annotations

import math

import numpy as np
import pytest

from fasterlmm_amd import expsum


@pytest.mark.parametrize("z_lo,z_hi", [(math.exp(-10), 3e4), (0.1, 3e4), (1e-3, 1e-3), (2.0, 5.0),
                                       (1e-12, 1e6)])
@pytest.mark.parametrize("eps", [1e-8, 1e-10, 1e-12])
def test_sum_matches_reciprocal(z_lo, z_hi, eps):
    t, lam = expsum.weight_nodes(z_lo, z_hi, eps)
    assert np.all(t > 0) and np.all(lam > 0)  # positive terms: no cancellation in the factored product
    z = np.geomspace(z_lo, z_hi, 20001) if z_hi > z_lo else np.array([z_lo])
    approx = (lam ** 2 * np.exp(-np.outer(z, t))).sum(axis=1)
    rel = np.abs(approx * z - 1.0)
    assert rel.max() <= 1.5 * eps, rel.max()


def test_term_count_for_the_scan_range():
    # log delta in [-10, 10] and eigenvalues up to 8,000: the sum fits one 128-entry K padding
    t, _ = expsum.weight_nodes(*expsum.z_range(0.0, 8000.0, -10.0, 10.0))
    assert t.size <= 128


def test_factored_sws_equals_exact_sum():
    """sum_k US_k^2 / (s_k + delta_p) == sum_r C[m, r] B[p, r] with C = US^2 A, for a singular spectrum and deltas
    across the grid (FP64 throughout: only the sum's own error remains)."""
    rng = np.random.default_rng(3)
    s = np.sort(rng.exponential(2.0, 300))
    s[:20] = 0.0
    s[-1] = 900.0
    US2 = rng.standard_normal((40, s.size)) ** 2
    log_delta = np.linspace(-10, 10, 57)
    t, lam = expsum.weight_nodes(*expsum.z_range(s.min(), s.max(), log_delta.min(), log_delta.max()))
    A = lam[None, :] * np.exp(-np.outer(s, t))
    B = lam[None, :] * np.exp(-np.outer(np.exp(log_delta), t))
    exact = US2 @ (1.0 / (s[:, None] + np.exp(log_delta)[None, :]))
    fact = (US2 @ A) @ B.T
    assert np.abs(fact / exact - 1).max() <= 1.5 * expsum.EPS


def test_bad_ranges_raise():
    for a, b in [(0.0, 1.0), (2.0, 1.0), (1.0, math.inf)]:
        with pytest.raises(ValueError):
            expsum.weight_nodes(a, b)


def test_engine_factored_weights_map():
    """engine.factored_weights (the SNP-side map A and the nodes) on CPU tensors: A is zero beyond Neff and R, the
    factored sWs matches the exact one, and no factoring is offered when the sum would not be shorter than K."""
    import torch
    from fasterlmm_amd import engine
    rng = np.random.default_rng(11)
    s = torch.from_numpy(np.sort(rng.exponential(1.5, 300)))
    s[:7] = 0.0
    rng_s = (float(s.min()), float(s.max()))
    fac = engine.factored_weights(s, rng_s, 384, -10.0, 10.0)
    assert fac is not None and fac.rpad == 128 and 0 < fac.R <= 128
    A = fac.A.numpy()
    assert A.shape == (384, 128) and not A[300:].any() and not A[:, fac.R:].any()
    log_delta = np.linspace(-10.0, 10.0, 41)
    B = fac.lam.numpy()[None, :] * np.exp(-np.outer(np.exp(log_delta), fac.t.numpy()))
    US2 = rng.standard_normal((25, 300)) ** 2
    exact = US2 @ (1.0 / (s.numpy()[:, None] + np.exp(log_delta)[None, :]))
    fact = (US2 @ A[:300, : fac.R]) @ B.T
    assert np.abs(fact / exact - 1).max() <= 1.5 * expsum.EPS
    assert engine.factored_weights(s, rng_s, 128, -10.0, 10.0) is None
