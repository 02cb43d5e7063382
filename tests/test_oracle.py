"""CPU tests that pin the FP64 oracle: against the reference implementation's golden vectors, and against
independent mathematics (dense GLS, OLS limits, closed forms, library routines) chosen so that a dropped term, a
wrong sign or index, or a transposed operand fails at least one of them."""
from __future__ import annotations

import numpy as np
import pytest
from scipy import optimize, stats

from conftest import golden
from oracle import lmm, perms


def _panel(N=40, M=120, P=3, C=2, seed=7):
    rng = np.random.default_rng(seed)
    Z = lmm.standardise_columns(rng.integers(0, 3, size=(N, M)).astype(np.float64))
    X = np.ones((N, C))
    X[:, 1:] = rng.normal(size=(N, C - 1))
    Y = rng.normal(size=(N, P)) + 0.5 * Z[:, :5].sum(axis=1, keepdims=True)
    return Z, X, Y


def _dense_gls(K, X, y, s, delta, *, dof):
    """Textbook GLS Wald test of y ~ X b + s beta with V = K + delta I (dense solves, no spectral tricks)."""
    V = K + delta * np.eye(K.shape[0])
    A = np.column_stack([X, s])
    Vi_A = np.linalg.solve(V, A)
    Vi_y = np.linalg.solve(V, y)
    G = A.T @ Vi_A
    coef = np.linalg.solve(G, A.T @ Vi_y)
    r = y - A @ coef
    sigma2 = r @ np.linalg.solve(V, r) / dof
    cov = sigma2 * np.linalg.inv(G)
    beta = coef[-1]
    se = np.sqrt(cov[-1, -1])
    return beta, se, beta * beta / cov[-1, -1]


# ---------------------------------------------------------------- reference golden vectors

@pytest.mark.parametrize("tag", ["int", "cov2", "full"])
def test_compat_scan_matches_reference_at_reference_delta(example, tag):
    """Per chromosome, with the reference's own refined log delta, the oracle's compat rotation + scan agree with
    the reference's LOCO outputs to ~1e-10, and its grid argmin is the reference's grid argmin."""
    ref = golden(f"compat_loco_{tag}.npz")
    Zs = lmm.standardise_columns(example["Z_raw"].astype(np.float64))
    chrom = example["chrom"]
    X = {"int": np.ones((Zs.shape[0], 1)), "cov2": example["X_cov2"], "full": example["X_full"]}[tag]
    Y = example["Y"][:, :3]
    for ci, c in enumerate(sorted(np.unique(chrom))):
        spec = lmm.compat_rotate(lmm.grm(Zs[:, chrom != c]), X, Y)
        out = lmm.snp_wald_scan_compat(spec, ref["log_delta"][ci], Zs[:, chrom == c])
        for k in ("f", "beta", "se", "sfve"):
            np.testing.assert_allclose(out[k], ref[k][chrom == c], rtol=1e-8, atol=1e-12)
        np.testing.assert_allclose(lmm.fit_delta_grid_compat(spec, refine=False), ref["log_delta_grid"][ci],
                                   atol=1e-13)


@pytest.mark.parametrize("tag", ["int", "cov2", "full"])
def test_compat_loco_end_to_end_matches_reference(example, tag):
    """End to end the oracle lands within the reference's own golden-section noise (~1e-6 in log delta)."""
    ref = golden(f"compat_loco_{tag}.npz")
    Zs = lmm.standardise_columns(example["Z_raw"].astype(np.float64))
    X = {"int": np.ones((Zs.shape[0], 1)), "cov2": example["X_cov2"], "full": example["X_full"]}[tag]
    out, lds = lmm.loco_scan_compat(Zs, X, example["Y"][:, :3], example["chrom"], return_log_delta=True)
    np.testing.assert_allclose(np.stack([lds[i] for i in range(len(lds))]), ref["log_delta"], atol=5e-6)
    df = Zs.shape[0] - X.shape[1] - 1
    lp = lambda F: stats.f.logsf(F, 1, df) / np.log(10)  # noqa: E731
    assert np.abs(lp(out["f"]) - lp(ref["f"])).max() < 1e-5
    np.testing.assert_allclose(out["max_F"], ref["max_F"], rtol=1e-6)
    np.testing.assert_allclose(out["nullh2"], ref["nullh2"], rtol=1e-6)


def test_plain_loco_matches_reference(example):
    ref = golden("plain_loco_cov2.npz")
    Zs = lmm.standardise_columns(example["Z_raw"].astype(np.float64))
    chrom = example["chrom"]
    X = example["X_cov2"]
    Y = example["Y"][:, :3]
    out = lmm.loco_scan(Zs, X, Y, chrom)
    df = Zs.shape[0] - X.shape[1] - 1
    lp = lambda F: stats.f.logsf(F, 1, df) / np.log(10)  # noqa: E731
    assert np.abs(lp(out["f"]) - lp(ref["f"])).max() < 1e-5
    for ci, c in enumerate(sorted(np.unique(chrom))[:4]):
        spec = lmm.rotate(lmm.grm(Zs[:, chrom != c]), X, Y)
        np.testing.assert_allclose(lmm.fit_delta_grid(spec), ref["log_delta_ml"][ci], atol=5e-6)
        np.testing.assert_allclose(lmm.fit_delta_grid(spec, reml=True), ref["log_delta_reml"][ci], atol=5e-6)
        o = lmm.snp_wald_scan(spec, ref["log_delta_ml"][ci], spec["U"].T @ Zs[:, chrom == c])
        np.testing.assert_allclose(o["f"], ref["f"][chrom == c], rtol=1e-8, atol=1e-12)


def test_perm_threshold_matches_reference(example):
    for name, idx, n_perm, loco in (("perm_loco_cov2.npz", [0, 1], 4, True), ("perm_single_cov2.npz", [3], 3, False)):
        ref = golden(name)
        res, pmf = perms.perm_threshold(example["Z_raw"].astype(np.float64), example["X_cov2"], example["Y"],
                                        example["chrom"], idx, n_perm=n_perm, loco=loco)
        np.testing.assert_allclose(pmf, ref["perm_max_F"], rtol=1e-6)
        np.testing.assert_allclose(res["f"], ref["f"], rtol=1e-3, atol=1e-6)


# ---------------------------------------------------------------- independent mathematics

def test_standardise_closed_form():
    a = np.array([[1.0, 10.0, 2.0], [np.nan, 20.0, 2.0], [3.0, np.nan, 2.0], [5.0, 40.0, 2.0], [7.0, 50.0, 2.0]])
    out = lmm.standardise_columns(a)
    x = np.array([1.0, 3.0, 5.0, 7.0])
    np.testing.assert_allclose(out[[0, 2, 3, 4], 0], (x - 4.0) / np.sqrt(5.0))  # ddof 0
    assert out[1, 0] == 0.0 and out[2, 1] == 0.0
    assert np.all(out[:, 2] == 0.0)


def test_grm_brute_force():
    rng = np.random.default_rng(1)
    Z = rng.normal(size=(6, 9))
    K = lmm.grm(Z)
    for i in range(6):
        for j in range(6):
            assert abs(K[i, j] - sum(Z[i, m] * Z[j, m] for m in range(9)) / 9) < 1e-14


def test_compat_scan_equals_dense_gls():
    """Compat F / beta / se at a given delta equal the textbook GLS Wald test with V = K + delta I and the
    covariates X as nuisance (dof N - C - 1): the projected basis carries the GLS Schur complement exactly."""
    Z, X, Y = _panel(N=36, M=30, P=2, C=3)
    K = lmm.grm(Z[:, 10:])
    spec = lmm.compat_rotate(K, X, Y)
    ld = np.array([-0.7, 1.3])
    out = lmm.snp_wald_scan_compat(spec, ld, Z[:, :10])
    for p in range(2):
        for m in range(10):
            b, se, F = _dense_gls(K, X, Y[:, p], Z[:, m], np.exp(ld[p]), dof=36 - 3 - 1)
            assert abs(out["beta"][m, p] - b) <= 1e-9 * max(1.0, abs(b))
            assert abs(out["se"][m, p] - se) <= 1e-9 * se
            assert abs(out["f"][m, p] - F) <= 1e-8 * max(1.0, F)


def test_plain_scan_equals_dense_gls():
    """Plain-path F / beta at delta equal dense GLS with V = K + (delta + jitter) I (core.py:89-102 jitter)."""
    Z, X, Y = _panel(N=30, M=25, P=2, C=2, seed=11)
    K = lmm.grm(Z[:, 8:])
    spec = lmm.rotate(K, X, Y)
    ld = np.array([0.4, -1.1])
    out = lmm.snp_wald_scan(spec, ld, spec["U"].T @ Z[:, :8])
    for p in range(2):
        for m in range(8):
            b, se, F = _dense_gls(K, X, Y[:, p], Z[:, m], np.exp(ld[p]) + 1e-6, dof=30 - 2 - 1)
            assert abs(out["beta"][m, p] - b) <= 1e-8 * max(1.0, abs(b))
            assert abs(out["f"][m, p] - F) <= 1e-7 * max(1.0, F)


def test_compat_loss_equals_restricted_likelihood():
    """Compat loss = Neff log(2 pi sigma2) + Neff + log det(Q^T (K + delta I) Q) with Q an orthonormal basis of
    null(X^T): the REML-type likelihood of the projected data, by dense slogdet."""
    Z, X, Y = _panel(N=30, M=60, P=2, C=2, seed=5)
    K = lmm.grm(Z)
    spec = lmm.compat_rotate(K, X, Y)
    Q = np.linalg.svd(np.eye(30) - X @ np.linalg.pinv(X))[0][:, :28]
    for ld in (-2.0, 0.3, 3.0):
        V = Q.T @ (K + np.exp(ld) * np.eye(30)) @ Q
        for p in range(2):
            yq = Q.T @ Y[:, p]
            rWr = yq @ np.linalg.solve(V, yq)
            want = 28 * (np.log(2 * np.pi * rWr / 28) + 1) + np.linalg.slogdet(V)[1]
            got = lmm.profile_loss_compat(np.array([ld]), spec["s"], spec["UY"])[0, p]
            assert abs(got - want) < 1e-9 * abs(want)


def test_plain_loss_equals_dense_ml_and_reml():
    Z, X, Y = _panel(N=28, M=50, P=1, C=2, seed=9)
    K = lmm.grm(Z)
    spec = lmm.rotate(K, X, Y)
    ld = 0.8
    V = K + (np.exp(ld) + 1e-6) * np.eye(28)
    y = Y[:, 0]
    Vi = np.linalg.inv(V)
    A = X.T @ Vi @ X
    b = np.linalg.solve(A, X.T @ Vi @ y)
    r = y - X @ b
    rWr = r @ Vi @ r
    ml = 28 * (np.log(2 * np.pi * rWr / 28) + 1) + np.linalg.slogdet(V)[1]
    reml = 26 * (np.log(2 * np.pi * rWr / 26) + 1) + np.linalg.slogdet(V)[1] + np.linalg.slogdet(A)[1]
    assert abs(lmm.profile_loss(np.array([ld]), spec["s"], spec["X_rot"], spec["Y_rot"])[0, 0] - ml) < 1e-8
    assert abs(lmm.profile_loss(np.array([ld]), spec["s"], spec["X_rot"], spec["Y_rot"], reml=True)[0, 0] - reml) < 1e-8
    assert abs(lmm.nll_eval(np.exp(ld), spec["s"], spec["X_rot"], y @ spec["U"]) - ml) < 1e-8


def test_ols_limit_large_delta():
    """delta -> infinity turns the GLS scan into OLS of y on [X, s]: F -> t^2 of numpy lstsq."""
    Z, X, Y = _panel(N=40, M=20, P=1, C=2, seed=3)
    spec = lmm.compat_rotate(lmm.grm(Z[:, 5:]), X, Y)
    out = lmm.snp_wald_scan_compat(spec, np.array([40.0]), Z[:, :5])
    for m in range(5):
        A = np.column_stack([X, Z[:, m]])
        coef, rss, *_ = np.linalg.lstsq(A, Y[:, 0], rcond=None)
        sig2 = rss[0] / (40 - 3)
        var_b = sig2 * np.linalg.inv(A.T @ A)[-1, -1]
        assert abs(out["beta"][m, 0] - coef[-1]) < 1e-9
        assert abs(out["f"][m, 0] - coef[-1] ** 2 / var_b) < 1e-7


def test_delta_search_grid_and_refinement():
    Z, X, Y = _panel(N=40, M=80, P=4, C=2, seed=21)
    spec = lmm.compat_rotate(lmm.grm(Z), X, Y)
    grid = np.linspace(-10, 10, 64)
    g = lmm.fit_delta_grid_compat(spec, refine=False)
    assert all(np.isclose(grid, v).any() for v in g)
    loss = lmm.profile_loss_compat(grid, spec["s"], spec["UY"])
    np.testing.assert_array_equal(g, grid[np.argmin(loss, axis=0)])
    ld = lmm.fit_delta_grid_compat(spec)
    for p in range(4):
        i = int(np.argmin(loss[:, p]))
        a, b = grid[max(i - 1, 0)], grid[min(i + 1, 63)]
        f = lambda t: lmm.profile_loss_compat_per_pheno(np.array([t]), spec["s"], spec["UY"][:, p:p + 1])[0]  # noqa
        best = optimize.minimize_scalar(f, bounds=(a, b), method="bounded", options={"xatol": 1e-10}).x
        assert abs(ld[p] - best) < 2e-5
        assert f(ld[p]) <= loss[i, p] + 1e-9
    # duplicate columns fit identically
    Y2 = np.repeat(Y[:, :1], 3, axis=1)
    spec2 = lmm.compat_rotate(lmm.grm(Z), X, Y2)
    ld2 = lmm.fit_delta_grid_compat(spec2)
    assert np.all(ld2 == ld2[0])


def test_compat_basis_invariants():
    Z, X, Y = _panel(N=32, M=90, P=3, C=4, seed=2)
    K = lmm.grm(Z)
    spec = lmm.compat_rotate(K, X, Y)
    assert spec["s"].shape == (28,)
    np.testing.assert_allclose(spec["U_eff"].T @ spec["U_eff"], np.eye(28), atol=1e-10)
    np.testing.assert_allclose(spec["U_eff"].T @ X, 0.0, atol=1e-10)
    assert np.all(np.diff(spec["s"]) >= -1e-12) and np.all(spec["s"] >= 0)
    Yr = Y - X @ (np.linalg.pinv(X) @ Y)
    np.testing.assert_allclose(spec["UY"], spec["U_eff"].T @ Yr, atol=1e-10)


def test_max_f_and_n_real_split():
    Z, X, Y = _panel(N=34, M=60, P=5, C=2, seed=4)
    spec = lmm.compat_rotate(lmm.grm(Z[:, 20:]), X, Y)
    ld = lmm.fit_delta_grid_compat(spec)
    full = lmm.snp_wald_scan_compat(spec, ld, Z[:, :20])
    part = lmm.snp_wald_scan_compat(spec, ld, Z[:, :20], n_real=2)
    np.testing.assert_array_equal(full["max_F"], full["f"].max(axis=0))
    np.testing.assert_array_equal(part["f"], full["f"][:, :2])
    np.testing.assert_array_equal(part["max_F"], full["max_F"])


def test_perm_orders_match_independent_regather():
    """Packing keyed per (seed, column) matches a per-column re-gather with default_rng([seed, p])
    (the scheme of the reference's tests/test_cli_extreme.py:48-66)."""
    rng = np.random.default_rng(123)
    N, n_perm, idx = 25, 6, [4, 0, 9]
    y = rng.standard_normal((N, 3))
    packed = perms.pack_perm_columns(y, perms.perm_orders(19930909, idx, N, n_perm))
    np.testing.assert_array_equal(packed[:, :3], y)
    for b, p in enumerate(idx):
        o = np.random.default_rng([19930909, p]).random((n_perm, N)).argsort(axis=1).T
        for j in range(n_perm):
            np.testing.assert_array_equal(packed[:, 3 + b * n_perm + j], y[o[:, j], b])


def test_f_pvalue_is_two_sided_t():
    import mpmath
    for F, df in ((0.5, 10), (4.0, 146), (30.0, 995), (400.0, 995)):
        p = perms.f_pvalue(F, df)
        t = np.sqrt(F)
        assert abs(p - 2 * stats.t.sf(t, df)) <= 1e-12 * max(p, 1e-300) + 1e-300
        x = df / (df + F)
        exact = float(mpmath.betainc(df / 2.0, 0.5, 0, x, regularized=True))
        assert abs(p - exact) <= 1e-10 * exact
        assert abs(perms.neglog10_pvalue(F, df) + np.log10(exact)) < 1e-9


def test_group_by_na():
    Y = np.arange(40.0).reshape(10, 4)
    Y[[0, 1], 0] = np.nan
    Y[[0, 1], 1] = np.nan
    Y[2, 2] = np.nan
    groups = {tuple(c): k for k, c in lmm.group_by_na(Y, [0, 1, 2, 3])}
    assert groups[(0, 1)] == list(range(2, 10))
    assert groups[(2,)] == [i for i in range(10) if i != 2]
    assert groups[(3,)] == list(range(10))


def test_loco_requires_kinship_pair():
    Z, X, Y = _panel()
    with pytest.raises(ValueError):
        lmm.loco_scan_compat(Z, X, Y, np.repeat([1, 2, 3], 40), Z_kin=Z)
