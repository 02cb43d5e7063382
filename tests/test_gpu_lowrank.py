"""GPU parity of the low-rank path (lowrank / extreme / extreme_scan) against the FP64 low-rank oracle.

Bars: |d log10 p| <= 1e-3 and |d beta| <= 1e-5 max(|beta|, se) (5e-5 for split-BF16, as in the compat tests)
per test; log delta to 1e-5 against the oracle's golden section; max F to 1e-4.  The scan kernel's second phase
(the off-rank cross term over K = N with the second product stopped at the on-rank extent) is exercised by
every scan here: panels span several SNP tiles (128) and column tiles (256) with ragged edges.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import bed as obed, lmm, lowrank as olr
from test_gpu_compat import BETA_BAR, check_scan

pytestmark = pytest.mark.gpu

DEV = "cuda"


def panel(N=300, Mkin=120, Mtest=300, P=24, n_perm=11, Dcov=3, seed=19930909):
    """Kinship markers G, test SNPs Z (standardised), X = [1 | covariates], Y = reals + row shuffles."""
    rng = np.random.default_rng(seed)
    G = lmm.standardise_columns(rng.binomial(2, 0.3, size=(N, Mkin)).astype(np.float64))
    Zraw = rng.binomial(2, 0.25, size=(N, Mtest)).astype(np.float64)
    Z = lmm.standardise_columns(Zraw)
    X = np.concatenate([np.ones((N, 1)), rng.standard_normal((N, Dcov - 1))], axis=1)
    Yr = G @ rng.standard_normal((Mkin, P)) / np.sqrt(Mkin) + 0.5 * Z[:, [0]] * rng.standard_normal((1, P))
    Yr += rng.standard_normal((N, P)) + X @ rng.standard_normal((Dcov, P))
    Y = np.concatenate([Yr] + [Yr[rng.permutation(N)] for _ in range(n_perm)], axis=1)
    return G, Z, Zraw, X, Y


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.mark.parametrize("precision", ["split_fp16", "split_bf16", "fp32"])
def test_single_k_lowrank_vs_oracle(precision):
    from fasterlmm_amd import engine, lowrank
    G, Z, _, X, Y = panel()
    n_real = 24
    spec_o = olr.lowrank_rotate(G, X, Y)
    ld_o = olr.fit_delta_grid_lowrank(spec_o)
    ref = olr.snp_wald_scan_lowrank(spec_o, ld_o, Z, n_real=n_real)
    with engine.configure(precision=precision):
        spec = lowrank.lowrank_rotate(t(G), t(X), t(Y))
        np.testing.assert_allclose(spec.s.cpu().numpy(), spec_o["s"], rtol=1e-10, atol=1e-12)
        ld_grid = lowrank.fit_delta_grid_lowrank(spec, refine=False).cpu().numpy()
        ld = lowrank.fit_delta_grid_lowrank(spec)
        res = lowrank.snp_wald_scan_lowrank(spec, ld, t(Z), n_real=n_real)
    # delta: the FP64 grid argmin is exact; the refined log delta matches the oracle's golden section where the
    # likelihood is curved and, everywhere, reaches a -2 log-likelihood within 1e-4 of it (on flat likelihoods,
    # h2 near 0, the refinement brackets the FP32 grid argmin, which can sit a node away from the FP64 one)
    np.testing.assert_allclose(ld_grid, olr.fit_delta_grid_lowrank(spec_o, refine=False), rtol=0, atol=1e-12)
    yy_off = (spec_o["UUY"] ** 2).sum(axis=0)
    loss = lambda v: olr.profile_loss_lowrank_per_pheno(v, spec_o["s"], spec_o["UY"], yy_off, spec_o["Neff"])  # noqa
    assert (loss(ld.cpu().numpy()) - loss(ld_o)).max() <= 1e-4
    curved = np.abs(ld_o) < 5.0
    assert np.abs(ld.cpu().numpy() - ld_o)[curved].max() <= 1e-5
    df = spec_o["Neff"] - 1
    check_scan(res, ref, df, beta_tol=BETA_BAR[precision], what=f"lowrank {precision}")
    np.testing.assert_allclose(res.max_F.cpu().numpy(), ref["max_F"], rtol=1e-4)
    np.testing.assert_allclose(res.nullh2.cpu().numpy(), ref["nullh2"], rtol=1e-5)


def test_scan_at_oracle_delta_and_block_invariance():
    """At the oracle's own log delta the scan alone is checked, and SNP blocks of 128 give the same results as one
    block (each SNP tile is computed identically wherever its block starts; the FP64 re-scoring of large-F tests
    may round its sums in another order)."""
    from fasterlmm_amd import lowrank
    G, Z, _, X, Y = panel(seed=7)
    spec_o = olr.lowrank_rotate(G, X, Y)
    ld_o = olr.fit_delta_grid_lowrank(spec_o)
    ref = olr.snp_wald_scan_lowrank(spec_o, ld_o, Z, n_real=5)
    spec = lowrank.lowrank_rotate(t(G), t(X), t(Y))
    a = lowrank.snp_wald_scan_lowrank(spec, t(ld_o), t(Z), n_real=5)
    b = lowrank.snp_wald_scan_lowrank(spec, t(ld_o), t(Z), n_real=5, snp_chunk=128)
    check_scan(a, ref, spec_o["Neff"] - 1, what="lowrank at oracle delta")
    for k in ("f", "beta", "se", "sfve", "nullh2", "max_F"):
        torch.testing.assert_close(getattr(a, k), getattr(b, k), rtol=1e-12, atol=0, msg=k)


def test_large_f_tests_are_rescored_in_fp64():
    """Tests with F above the re-scoring threshold carry FP64 sums: they match the oracle to ~1e-10, far inside
    the bars the MFMA sums meet."""
    from fasterlmm_amd import lowrank
    G, Z, _, X, Y = panel(seed=17)
    Y = Y.copy()
    Y[:, :4] += 3.0 * Z[:, [5]]  # plant strong hits so some real tests pass the threshold
    spec_o = olr.lowrank_rotate(G, X, Y)
    ld_o = olr.fit_delta_grid_lowrank(spec_o)
    ref = olr.snp_wald_scan_lowrank(spec_o, ld_o, Z, n_real=8)
    spec = lowrank.lowrank_rotate(t(G), t(X), t(Y))
    res = lowrank.snp_wald_scan_lowrank(spec, t(ld_o), t(Z), n_real=8)
    big = ref["f"] > 2 * lowrank.RESCORE_F
    assert big.sum() >= 4
    np.testing.assert_allclose(res.f.cpu().numpy()[big], ref["f"][big], rtol=1e-9)
    np.testing.assert_allclose(res.beta.cpu().numpy()[big], ref["beta"][big], rtol=1e-9)
    hit = ref["max_F"][:8] > 2 * lowrank.RESCORE_F  # columns whose maximum is a re-scored test
    assert hit.sum() >= 2
    np.testing.assert_allclose(res.max_F.cpu().numpy()[:8][hit], ref["max_F"][:8][hit], rtol=1e-9)
    np.testing.assert_allclose(res.max_F.cpu().numpy()[:8], ref["max_F"][:8], rtol=1e-5)


def test_loco_lowrank_vs_oracle():
    from fasterlmm_amd import lowrank
    G, Z, _, X, Y = panel(N=240, Mkin=90, Mtest=260, P=6, n_perm=20, seed=3)
    g_chrom = np.repeat([1, 2, 3], [30, 30, 30])
    z_chrom = np.repeat([1, 2, 3], [100, 90, 70])
    ref = olr.loco_scan_lowrank(G, g_chrom, Z, z_chrom, X, Y, n_real=6)
    res = lowrank.loco_scan_lowrank(t(G), g_chrom, t(Z), z_chrom, t(X), t(Y), n_real=6)
    check_scan(res, ref, 240 - 3 - 1, what="loco lowrank")
    np.testing.assert_allclose(res.max_F.cpu().numpy(), ref["max_F"], rtol=1e-4)


def test_streamed_equals_resident(tmp_path):
    """The streamed engine (decode from .bed per block) and the resident one (FP64 standardised matrix in HBM)
    give identical bits; the int8 resident form matches to rounding; all three match the oracle."""
    from fasterlmm_amd import bed, extreme_scan
    G, _, Zraw, X, Y = panel(N=200, Mkin=60, Mtest=330, P=4, n_perm=30, seed=11)
    n, m = Zraw.shape
    chrom = np.repeat([1, 2], [170, 160])
    iid = [f"s{i}" for i in range(n)]
    prefix = tmp_path / "g"
    (tmp_path / "g.bed").write_bytes(obed.encode(Zraw).tobytes())
    (tmp_path / "g.fam").write_text("".join(f"F {s} 0 0 0 -9\n" for s in iid))
    (tmp_path / "g.bim").write_text("".join(f"{int(c)}\trs{j}\t0\t{1000 + j}\tA\tG\n" for j, c in enumerate(chrom)))
    h = bed.open_aligned_bed(prefix, iid)
    g_chrom = np.repeat([1, 2], [30, 30])
    a = extreme_scan.loco_scan_streamed(h, t(G), g_chrom, t(X), t(Y), n_real=4, block_size=100)
    Zs = bed.read_all_standardised(h, dtype=torch.float64)
    b = extreme_scan.loco_scan_resident(Zs, np.asarray(h.chrom), t(G), g_chrom, t(X), t(Y), n_real=4,
                                        block_size=100)
    for k in ("f", "beta", "se", "sfve", "nullh2", "max_F"):
        assert torch.equal(getattr(a, k), getattr(b, k)), k
    c = extreme_scan.loco_scan_resident(t(Zraw.astype(np.int8)), chrom, t(G), g_chrom, t(X), t(Y), n_real=4,
                                        block_size=64)
    for k in ("f", "beta", "max_F"):  # the two standardisations differ in the last FP64 bit
        torch.testing.assert_close(getattr(c, k), getattr(a, k), rtol=1e-6, atol=1e-9)
    ref = olr.loco_scan_lowrank(G, g_chrom, lmm.standardise_columns(Zraw), chrom, X, Y, n_real=4)
    check_scan(a, ref, n - 3 - 1, what="streamed")
    check_scan(c, ref, n - 3 - 1, what="resident int8")


def test_randomized_eigh_and_rank_truncated_basis():
    """The range finder recovers the leading spectrum across a gap (against numpy eigh); a rank-truncated basis
    is orthonormal, orthogonal to X and spans the leading directions of the regressed factor; with rank >= the
    factor's rank it equals the exact basis and the scan equals the exact low-rank scan."""
    from fasterlmm_amd import extreme, lowrank
    rng = np.random.default_rng(5)
    N, M = 400, 900
    Q, _ = np.linalg.qr(rng.standard_normal((N, N)))
    sv = np.concatenate([np.linspace(40, 12, 16), np.full(N - 16, 0.4)])
    Zk = Q @ np.diag(sv) @ np.linalg.qr(rng.standard_normal((M, N)))[0].T
    s, U = extreme.randomized_eigh(t(Zk), rank=16, n_iter=3)
    ev = np.linalg.eigvalsh(Zk @ Zk.T / M)[::-1][:16]
    np.testing.assert_allclose(s.cpu().numpy(), ev, rtol=1e-8)
    G, Z, _, X, Y = panel(N=300, Mkin=60, Mtest=200, P=3, n_perm=5, seed=13)
    b = lowrank.lowrank_basis(t(G), t(X), rank=20, n_oversample=40)  # probe spans the factor's whole range
    Ub = b.U_eff.cpu().numpy()
    assert Ub.shape == (300, 20)
    np.testing.assert_allclose(Ub.T @ Ub, np.eye(20), atol=1e-10)
    assert np.abs(X.T @ Ub).max() < 1e-9
    ex = olr.lowrank_basis(G, X)
    np.testing.assert_allclose(b.s.cpu().numpy(), ex["s"][-20:], rtol=1e-8)
    # rank above the factor's rank: exact branch
    full = lowrank.single_k_scan_lowrank(t(G), t(Z), t(X), t(Y), n_real=3, rank=10_000)
    ref = olr.single_k_scan_lowrank(G, Z, X, Y, n_real=3)
    check_scan(full, ref, 300 - 3 - 1, what="rank >= factor rank")
