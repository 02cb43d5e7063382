"""Pins of the low-rank oracle (oracle/lowrank.py) by facts independent of its own formulas.

* the basis spans the X-regressed kinship factor and diagonalises the projected kinship;
* the low-rank likelihood and scan equal the full-rank compat ones (oracle.lmm, itself pinned by the reference's
  golden vectors) on the same kinship K = G G^T / k_in -- the reference's own cross-check
  (tests/test_lowrank.py:124-151) -- at a common log delta;
* LOCO is the per-chromosome single-kinship scan (tests/test_lowrank.py:154-187);
* the randomised range finder is exact on low-rank matrices and recovers the leading spectrum across a gap.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import lmm, lowrank as olr


def toy(N=80, Mkin=50, Mtest=30, P=5, Dcov=2, seed=19930909):
    """Kinship factor G, test SNPs Z (standardised), X = [1 | covariates], Y with polygenic + planted effects
    (the reference's toy shape, tests/test_lowrank.py:20-32)."""
    rng = np.random.default_rng(seed)
    G = lmm.standardise_columns(rng.binomial(2, 0.3, size=(N, Mkin)).astype(np.float64))
    Z = lmm.standardise_columns(rng.binomial(2, 0.25, size=(N, Mtest)).astype(np.float64))
    X = np.concatenate([np.ones((N, 1)), rng.standard_normal((N, Dcov))], axis=1)
    Y = G @ rng.standard_normal((Mkin, P)) / np.sqrt(Mkin) + 0.4 * Z[:, [0]] * rng.standard_normal((1, P))
    Y += rng.standard_normal((N, P)) + X @ rng.standard_normal((X.shape[1], P))
    return G, Z, X, Y


def test_basis_spans_projected_kinship():
    G, Z, X, Y = toy()
    b = olr.lowrank_basis(G, X)
    U, s = b["U_eff"], b["s"]
    k = s.shape[0]
    assert k <= G.shape[1]
    np.testing.assert_allclose(U.T @ U, np.eye(k), atol=1e-12)
    assert np.abs(X.T @ U).max() < 1e-10
    Iperp = np.eye(G.shape[0]) - X @ lmm.pinv(X)
    Kp = Iperp @ G @ G.T @ Iperp / G.shape[1]
    np.testing.assert_allclose(U @ np.diag(s) @ U.T, Kp, atol=1e-12)
    # the k >= N branch (eigh of the N x N kernel) keeps the same spectrum
    Gw = np.concatenate([G, G[:, :40]], axis=1)[:, : G.shape[0] + 5]
    bw = olr.lowrank_basis(Gw, X)
    Gw_ = Iperp @ Gw
    sv2 = np.linalg.svd(Gw_, compute_uv=False) ** 2 / Gw.shape[1]
    np.testing.assert_allclose(np.sort(bw["s"])[::-1], sv2[: bw["s"].shape[0]], rtol=1e-10, atol=1e-12)


def test_rotate_splits_on_and_off_rank():
    G, Z, X, Y = toy()
    spec = olr.lowrank_rotate(G, X, Y)
    Y_r = Y - X @ (lmm.pinv(X) @ Y)
    np.testing.assert_allclose(spec["U_eff"] @ spec["UY"] + spec["UUY"], Y_r, atol=1e-12)
    assert np.abs(spec["U_eff"].T @ spec["UUY"]).max() < 1e-10


def _compat_twin(G, X, Y):
    return lmm.compat_rotate(G @ G.T / G.shape[1], X, Y)


def test_loss_equals_full_rank_compat():
    G, Z, X, Y = toy()
    spec = olr.lowrank_rotate(G, X, Y)
    cmp = _compat_twin(G, X, Y)
    yy_off = (spec["UUY"] ** 2).sum(axis=0)
    grid = np.linspace(-6, 6, 25)
    a = olr.profile_loss_lowrank(grid, spec["s"], spec["UY"], yy_off, spec["Neff"])
    b = lmm.profile_loss_compat(grid, cmp["s"], cmp["UY"])
    np.testing.assert_allclose(a, b, rtol=1e-9)
    ld = np.linspace(-2, 2, Y.shape[1])
    np.testing.assert_allclose(olr.profile_loss_lowrank_per_pheno(ld, spec["s"], spec["UY"], yy_off, spec["Neff"]),
                               lmm.profile_loss_compat_per_pheno(ld, cmp["s"], cmp["UY"]), rtol=1e-9)


def test_scan_equals_full_rank_compat_at_common_delta():
    G, Z, X, Y = toy()
    spec = olr.lowrank_rotate(G, X, Y)
    cmp = _compat_twin(G, X, Y)
    ld = olr.fit_delta_grid_lowrank(spec)
    a = olr.snp_wald_scan_lowrank(spec, ld, Z, n_real=3)
    b = lmm.snp_wald_scan_compat(cmp, ld, Z, n_real=3)
    for k in ("f", "beta", "se", "sfve", "nullh2", "max_F"):
        np.testing.assert_allclose(a[k], b[k], rtol=1e-8, err_msg=k)
    # and the two delta fits land together (separate golden sections: the reference's ~1e-5 bar)
    np.testing.assert_allclose(ld, lmm.fit_delta_grid_compat(cmp), atol=1e-5)


def test_loco_is_per_chromosome_single_k():
    G, Z, X, Y = toy()
    g_chrom = np.repeat([1, 2, 3], [17, 17, 16])
    z_chrom = np.repeat([1, 2, 3], [10, 10, 10])
    out = olr.loco_scan_lowrank(G, g_chrom, Z, z_chrom, X, Y, n_real=2)
    for c in (1, 2, 3):
        r = olr.single_k_scan_lowrank(G[:, g_chrom != c], Z[:, z_chrom == c], X, Y, n_real=2)
        for k in ("f", "beta", "se", "sfve", "nullh2"):
            np.testing.assert_array_equal(out[k][z_chrom == c], r[k])


@pytest.mark.parametrize("n_iter", [0, 2])
def test_randomized_eigh_exact_on_low_rank(n_iter):
    rng = np.random.default_rng(3)
    N, M, r = 60, 200, 12
    Z = rng.standard_normal((N, r)) @ rng.standard_normal((r, M))  # rank r
    s, U = olr.randomized_eigh(Z, 8, rng.standard_normal((N, r + 4)), n_iter=n_iter)
    ev, V = np.linalg.eigh(Z @ Z.T / M)
    np.testing.assert_allclose(s, ev[::-1][:8], rtol=1e-10)
    np.testing.assert_allclose(np.abs(U.T @ V[:, ::-1][:, :8]), np.eye(8), atol=1e-8)


def test_randomized_eigh_recovers_leading_spectrum_across_gap():
    rng = np.random.default_rng(4)
    N, M = 80, 400
    Q, _ = np.linalg.qr(rng.standard_normal((N, N)))
    sv = np.concatenate([np.linspace(30, 10, 10), np.full(N - 10, 0.5)])
    Z = Q @ np.diag(sv) @ np.linalg.qr(rng.standard_normal((M, N)))[0].T
    s, U = olr.randomized_eigh(Z, 10, rng.standard_normal((N, 20)), n_iter=3)
    ev = np.linalg.eigvalsh(Z @ Z.T / M)[::-1][:10]
    np.testing.assert_allclose(s, ev, rtol=1e-6)
