"""GPU parity at the workloads' own sizes, against the FP64 oracle.

* N = 1,000 genome (the headline shape: M = 100,000 SNPs over 22 chromosomes, 6,484 RINT phenotypes, C = 4): one
  chromosome's 4,546 SNPs scanned with its LOCO kinship against every real phenotype, 12,968 permutation columns
  and planted columns whose delta sits at the grid's lower edge (h2 = 1), 100-point grid.  The oracle checks
  every real column on a SNP sample and the genome-wide maximum F of a sample of permutation columns over all
  of the chromosome's SNPs (the threshold's input).
* The delta search's global-memory path (neff > 1,024: N = 1,300), compat and plain-ML losses, Newton / golden /
  grid modes.
* Uneven missingness (N = 5,000, four NA groups keeping 3,000 - 5,000 strains, C = 6) through the grouped launch.
* Low-rank LOCO at N = 6,000 with a rank-600 kinship.
* The learned delta proposal at its evaluation size (N = 2,000, 1,000 phenotypes x (1 + 10 permutations)) against
  grid + Newton and the oracle.
"""
from __future__ import annotations

import time

import numpy as np
import pytest
import torch

from oracle import lmm, lowrank as olr, perms as operms
from test_gpu_compat import check_scan

pytestmark = pytest.mark.gpu
DEV = "cuda"


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def say(msg):
    print(msg, flush=True)


def _sub(spec, cols):
    """The oracle's compat spectrum restricted to some rotated columns (each column's algebra is independent)."""
    return {**spec, "UY": spec["UY"][:, cols]}


def test_headline_shape_chromosome_slice():
    from fasterlmm_amd import core, sim
    from fasterlmm_amd.io import standardise_columns
    t0 = time.perf_counter()
    pan = sim.make_panel(2)
    N, M = pan.G.shape
    chrom = pan.chrom
    c = 1
    on = chrom == c
    B = pan.Y.shape[1]
    n_perm = 2
    Zs = standardise_columns(T(pan.G))  # FP64 on the GPU, the scan's own standardisation
    Zs_h = Zs.cpu().numpy()
    # planted columns in the span of the three leading eigenvectors of chromosome 1's LOCO kinship (the population
    # structure axes, eigenvalues ~40 against a median ~0.9), no noise: the ML delta sits at the grid's lower
    # edge (h2 = 1); two ride as real columns, two in the permutation region (max-F-only tiles)
    rng = np.random.default_rng(11)
    K = lmm.grm(Zs_h[:, ~on])
    planted = np.linalg.eigh(K)[1][:, -3:] @ rng.standard_normal((3, 4))
    orders = operms.perm_orders(19930909, list(range(B)), N, n_perm)
    Yp = operms.pack_perm_columns(pan.Y, orders)  # [reals | b0 perms | b1 perms | ...]
    Y_all = np.concatenate([Yp[:, :B], planted[:, :2], Yp[:, B:], planted[:, 2:]], axis=1)
    n_real = B + 2
    P = Y_all.shape[1]
    say(f"headline slice: N={N} M={M} chromosome {c} ({int(on.sum())} SNPs) x {P} columns; panel "
        f"{time.perf_counter() - t0:.1f}s")
    res = core.loco_scan_compat(Zs[:, T(np.flatnonzero(on))], T(pan.X), T(Y_all), chrom[on], n_real=n_real,
                                Z_kin=Zs, chrom_kin=chrom, n_grid=100)
    torch.cuda.synchronize()
    f_gpu = res.f.cpu().numpy()
    say(f"GPU scan done {time.perf_counter() - t0:.1f}s")

    # oracle: every real column, 3,000 sampled permutation columns and the two planted permutation-region columns
    perm_cols = np.sort(np.concatenate([rng.choice(np.arange(n_real, P - 2), size=3000, replace=False),
                                        [P - 2, P - 1]]))
    cols = np.concatenate([np.arange(n_real), perm_cols])
    spec = lmm.compat_rotate(K, pan.X, Y_all[:, cols])
    ld = lmm.fit_delta_grid_compat(spec, n_grid=100)
    say(f"oracle delta done {time.perf_counter() - t0:.1f}s")
    S = Zs_h[:, on]
    # real columns: a 1% SNP sample plus the 50 SNPs with the largest GPU F
    top = np.argsort(f_gpu.max(axis=1))[-50:]
    snp = np.unique(np.concatenate([rng.choice(S.shape[1], size=S.shape[1] // 100, replace=False), top]))
    ref = lmm.snp_wald_scan_compat(_sub(spec, np.arange(n_real)), ld[:n_real], S[:, snp], n_real=n_real)
    df = N - pan.X.shape[1] - 1
    got = {k: getattr(res, k).cpu().numpy()[snp] for k in ("f", "beta", "se")}
    dlp, dbeta, drel = check_scan(got, ref, df, what="headline slice, real columns")
    # delta: the GPU's log delta (from nullh2) reaches the oracle's loss to within 1e-6 on every real column
    h2 = res.nullh2.cpu().numpy()[0]
    ld_gpu = np.log(1.0 / h2 - 1.0)
    loss = lambda v: lmm.profile_loss_compat_per_pheno(v, spec["s"], spec["UY"][:, :n_real])  # noqa: E731
    gap = (loss(ld_gpu) - loss(ld[:n_real])).max()
    assert gap <= 1e-6, gap
    assert ld[B] <= -9.0 and ld[B + 1] <= -9.0, ld[B:B + 2]  # the planted real columns sit at the lower edge
    # permutation columns: genome-wide max F over all of the chromosome's SNPs (the threshold's input)
    sel = np.arange(n_real, len(cols))
    ref_p = lmm.snp_wald_scan_compat(_sub(spec, sel), ld[sel], S, n_real=0)
    mf = res.max_F.cpu().numpy()[perm_cols]
    rel = np.abs(mf - ref_p["max_F"]) / ref_p["max_F"]
    assert ld[-2] <= -9.0 and ld[-1] <= -9.0, ld[-2:]  # the planted permutation-region columns too
    say(f"headline slice: max |d log10 p| {dlp:.2e}, max |d beta|/max(|beta|,se) {dbeta:.2e}, max |d beta|/|beta| "
        f"(|beta|>=se) {drel:.2e}, loss gap {gap:.1e}, perm max_F rel err {rel.max():.2e} (edge columns "
        f"{rel[-2:]}), {time.perf_counter() - t0:.1f}s")
    assert rel.max() <= 1e-4, (rel.max(), perm_cols[np.argmax(rel)])
    # real columns' max_F is their column maximum of f
    np.testing.assert_array_equal(res.max_F.cpu().numpy()[:n_real], f_gpu.max(axis=0))


def _panel_1300():
    from fasterlmm_amd import sim
    cfg = dict(N=1300, M=6000, n_chrom=3, P=40, C=3, maf=(0.05, 0.5), h2=(0.0, 0.95), rint=False)
    pan = sim.make_panel(cfg, seed=5)
    Zs = lmm.standardise_columns(pan.G.astype(np.float64))
    return pan, Zs


@pytest.mark.parametrize("mode", ["newton", "golden", "grid"])
def test_delta_search_global_memory_path(mode):
    """neff, nlam > 1,024: the refinement re-reads each column from memory every pass (no register copy).  Compat
    (nlam = neff) and plain ML (nlam = N, its log det over the full spectrum) against the oracle: grid nodes exact,
    refined log delta within 1e-4, the scan at the fitted delta to the scan bars on a SNP sample."""
    from fasterlmm_amd import core, engine
    pan, Zs = _panel_1300()
    on = pan.chrom == 2
    K = lmm.grm(Zs[:, ~on])
    Y = pan.Y
    refine = mode != "grid"
    S = Zs[:, np.flatnonzero(on)[::10]]
    df = Zs.shape[0] - pan.X.shape[1] - 1
    with engine.configure(delta="golden" if mode == "golden" else "newton"):
        spec = core.fastlmm_compat_rotate(T(K), T(pan.X), T(Y))
        assert spec.s.numel() > 1024
        ld = core.fit_delta_grid_compat(spec, refine=refine).cpu().numpy()
        want = lmm.compat_rotate(K, pan.X, Y)
        ld_o = lmm.fit_delta_grid_compat(want, refine=refine)
        res = core.snp_wald_scan_compat(spec, T(ld), T(S), n_real=Y.shape[1])
        ref = lmm.snp_wald_scan_compat(want, ld, S)
        check_scan(res, ref, Zs.shape[0] - pan.X.shape[1] - 1, what=f"compat neff>1024 {mode}")
        if refine:
            assert np.abs(ld - ld_o).max() <= 1e-4, np.abs(ld - ld_o).max()
        else:
            np.testing.assert_allclose(ld, ld_o, rtol=0, atol=1e-12)
        # plain GLS, ML loss: its log det runs over all N = 1,300 eigenvalues of K + 1e-6 I
        pspec = core.rotate(T(K), T(pan.X), T(Y))
        ldp = core.fit_delta_grid(pspec, refine=refine).cpu().numpy()
        pwant = lmm.rotate(K, pan.X, Y)
        ldp_o = lmm.fit_delta_grid(pwant, refine=refine)
        if refine:
            assert np.abs(ldp - ldp_o).max() <= 1e-4, np.abs(ldp - ldp_o).max()
        else:
            np.testing.assert_allclose(ldp, ldp_o, rtol=0, atol=1e-12)
        # each side rotates the SNPs by its own eigenvectors (defined up to a rotation in degenerate eigenspaces)
        res = core.snp_wald_scan(pspec, T(ldp), pspec.U.T @ T(S), n_real=Y.shape[1])
        ref = lmm.snp_wald_scan(pwant, ldp, pwant["U"].T @ S)
        check_scan(res, ref, df, what=f"plain ML neff>1024 {mode}")
    say(f"global-memory delta path {mode}: compat max |d log delta| {np.abs(ld - ld_o).max():.1e}, "
        f"plain {np.abs(ldp - ldp_o).max():.1e}")


def test_uneven_groups_at_workload_size():
    """Four missingness groups keeping 3,000 - 5,000 of N = 5,000 strains (C = 6), 8 phenotypes x (1 + 10
    permutations) each, through one grouped launch per SNP block; chromosome 1 scanned against its LOCO kinship.
    Each group's kinship has neff > 1,024, so its delta search takes the global-memory path."""
    from fasterlmm_amd import sim
    from fasterlmm_amd.io import align_arrays
    from fasterlmm_amd.perms import perm_threshold_groups
    cfg = dict(N=5000, M=16000, n_chrom=4, P=32, C=6, maf=(0.01, 0.5), h2=(0.0, 0.9), rint=True,
               na_groups=4, n_keep=(3000, 5000), sampler="bernoulli2")
    t0 = time.perf_counter()
    pan = sim.make_panel(cfg, seed=3)
    iid = [str(i) for i in range(cfg["N"])]
    data = align_arrays(iid, pan.G.astype(np.float64), iid, pan.Y, iid, pan.X[:, 1:], chrom=pan.chrom)
    n_perm = 10
    res = perm_threshold_groups(data, list(range(cfg["P"])), n_perm=n_perm, seed=19930909, test_chroms=[1])
    torch.cuda.synchronize()
    say(f"uneven groups: {len(res)} groups scanned in {time.perf_counter() - t0:.1f}s")
    assert len(res) == 4
    on = pan.chrom == 1
    X = pan.X
    rng = np.random.default_rng(2)
    snp = np.sort(rng.choice(np.flatnonzero(on), size=60, replace=False))
    for keep, cols, r, pmf in res:
        keep = np.asarray(keep)
        assert 3000 <= len(keep) <= 5000
        Zs = lmm.standardise_columns(pan.G[keep].astype(np.float64))
        y = pan.Y[keep][:, cols]
        orders = operms.perm_orders(19930909, cols, len(keep), n_perm)
        Yg = operms.pack_perm_columns(y, orders)
        spec = lmm.compat_rotate(lmm.grm(Zs[:, ~on]), X[keep], Yg)
        ld = lmm.fit_delta_grid_compat(spec)
        B = len(cols)
        ref = lmm.snp_wald_scan_compat(spec, ld, Zs[:, snp], n_real=B)
        got = {k: getattr(r, k).cpu().numpy()[snp] for k in ("f", "beta", "se")}
        df = len(keep) - X.shape[1] - 1
        dlp, dbeta, drel = check_scan(got, ref, df, what=f"group of {len(keep)} strains")
        full = lmm.snp_wald_scan_compat(spec, ld, Zs[:, on], n_real=0)
        rel = np.abs(pmf.cpu().numpy().reshape(-1) - full["max_F"][B:]) / full["max_F"][B:]
        assert rel.max() <= 1e-4, rel.max()
        say(f"group of {len(keep)}: max |d log10 p| {dlp:.2e}, d beta {dbeta:.2e}, rel {drel:.2e}, perm max_F "
            f"{rel.max():.2e} ({time.perf_counter() - t0:.1f}s)")


def test_lowrank_loco_at_reduced_cohort_size():
    """Low-rank LOCO on N = 6,000 with a rank-600 kinship (randomised range finder) of 3,000 markers over two
    chromosomes, C = 10, 40 phenotypes x (1 + 5 permutations): the GPU's basis and delta per chromosome, and the
    FP64 low-rank oracle's scan on a SNP sample with that basis and delta."""
    from fasterlmm_amd import engine, lowrank
    from fasterlmm_amd.io import standardise_columns
    rng = np.random.default_rng(13)
    N, Mk, Mt, C, B, n_perm, rank = 6000, 3000, 1200, 10, 40, 5, 600
    Gk = standardise_columns(T(rng.binomial(2, rng.uniform(0.05, 0.5, Mk), size=(N, Mk)).astype(np.float64)))
    Zt = standardise_columns(T(rng.binomial(2, rng.uniform(0.05, 0.5, Mt), size=(N, Mt)).astype(np.float64)))
    g_chrom = np.repeat([1, 2], Mk // 2)
    z_chrom = np.repeat([1, 2], Mt // 2)
    X = np.concatenate([np.ones((N, 1)), rng.standard_normal((N, C - 1))], axis=1)
    h2 = rng.uniform(0.1, 0.8, B)
    g = Gk.cpu().numpy() @ rng.standard_normal((Mk, B)) / np.sqrt(Mk)
    Yr = np.sqrt(h2) * g / g.std(0) + np.sqrt(1 - h2) * rng.standard_normal((N, B))
    Y = np.concatenate([Yr] + [Yr[rng.permutation(N)] for _ in range(n_perm)], axis=1)
    res = lowrank.loco_scan_lowrank(Gk, g_chrom, Zt, z_chrom, T(X), T(Y), n_real=B, rank=rank)
    torch.cuda.synchronize()
    df = N - C - 1
    for c in (1, 2):
        keep = torch.from_numpy((g_chrom != c).astype(np.float64)).to(DEV)
        b = lowrank._basis(Gk, T(X), keep, rank=rank)  # packs the same factor planes as the LOCO run: same bits
        UY, UUY = lowrank.rotate_phenos(b, T(Y))
        spec = lowrank.LowRankSpectrum(s=b.s, U_eff=b.U_eff, X=b.X, Xpinv=b.Xpinv, Neff=b.Neff, UY=UY, UUY=UUY)
        ld = lowrank.fit_delta_grid_lowrank(spec).cpu().numpy()
        assert b.s.numel() == rank
        ospec = {"s": b.s.cpu().numpy(), "U_eff": b.U_eff.cpu().numpy(), "X": X, "Xpinv": b.Xpinv.cpu().numpy(),
                 "Neff": b.Neff, "UY": UY.cpu().numpy(), "UUY": UUY.cpu().numpy()}
        # the GPU's delta is the oracle's golden-section optimum on the same basis (to the loss's flatness)
        yy_off = (ospec["UUY"] ** 2).sum(axis=0)
        loss = lambda v: olr.profile_loss_lowrank_per_pheno(v, ospec["s"], ospec["UY"], yy_off, b.Neff)  # noqa
        ld_o = olr.fit_delta_grid_lowrank(ospec)
        assert (loss(ld) - loss(ld_o)).max() <= 1e-6
        cols = np.flatnonzero(z_chrom == c)[::8]
        ref = olr.snp_wald_scan_lowrank(ospec, ld, Zt[:, T(cols)].cpu().numpy())
        got = {k: getattr(res, k).cpu().numpy()[cols] for k in ("f", "beta", "se")}
        dlp, dbeta, drel = check_scan(got, {k: ref[k][:, :B] for k in ("f", "beta", "se")}, df,
                                      what=f"low-rank chromosome {c}")
        say(f"low-rank N={N} rank {rank} chromosome {c}: max |d log10 p| {dlp:.2e}, d beta {dbeta:.2e}, "
            f"rel {drel:.2e}")


def test_learned_delta_at_evaluation_size():
    """The learned delta proposal (the shipped network) at N = 2,000: a Balding-Nichols panel (8,000 kinship
    markers, 200 test SNPs), C = 4, 1,000 phenotypes (h2 ~ U(0, 0.95)) x (1 + 10 permutations) = 11,000 columns,
    100-point grid.  Against grid + Newton: the same log delta (accepted proposals within 1e-6, fallback columns
    exactly).  Against the FP64 oracle on 100 real and 100 permutation columns: the oracle's golden-section optimum
    is no better than the learned one by more than 1e-6 in the loss, and the scan of the 200 SNPs at the learned
    delta meets the scan bars on the real columns."""
    from fasterlmm_amd import core, engine, sim
    from fasterlmm_amd.io import standardise_columns
    t0 = time.perf_counter()
    rng = np.random.default_rng(21)
    N, Mk, Mt, B, n_perm, C, n_grid = 2000, 8000, 200, 1000, 10, 4, 100
    Zs = standardise_columns(T(sim.balding_nichols(rng, N, Mk + Mt, fst=0.05, maf=(0.05, 0.5)).astype(np.float64)))
    Zk = Zs[:, :Mk]
    St = Zs[:, Mk:].cpu().numpy()
    K = ((Zk @ Zk.T) / Mk).cpu().numpy()
    X = np.concatenate([np.ones((N, 1)), rng.standard_normal((N, C - 1))], axis=1)
    h2 = rng.uniform(0.0, 0.95, B)
    g = Zk.cpu().numpy() @ rng.standard_normal((Mk, B))
    g = (g - g.mean(0)) / g.std(0)
    y = np.sqrt(h2) * g + np.sqrt(1.0 - h2) * rng.standard_normal((N, B))
    Y = np.concatenate([y] + [y[rng.permutation(N)] for _ in range(n_perm)], axis=1)
    P = Y.shape[1]
    fits = {}
    for mode in ("newton", "learned"):
        with engine.configure(delta=mode):
            spec = core.fastlmm_compat_rotate(T(K), T(X), T(Y))
            torch.cuda.synchronize()
            a = time.perf_counter()
            ld = core.fit_delta_grid_compat(spec, n_grid=n_grid)
            torch.cuda.synchronize()
            fits[mode] = (spec, ld, time.perf_counter() - a, getattr(spec._gpu, "last_fallback", None))
    spec, ld_l, t_l, fb = fits["learned"]
    ld_n, t_n = fits["newton"][1], fits["newton"][2]
    d = (ld_l - ld_n).abs().cpu().numpy()
    fb = fb.cpu().numpy()
    assert np.all(d[fb == 1] <= 1e-12), d[fb == 1].max()
    if (fb == 0).any():
        assert d[fb == 0].max() <= 1e-6, d[fb == 0].max()
    # the oracle on a column sample: real columns first (they carry per-test outputs), then permutation columns
    cols = np.concatenate([np.arange(100), B + rng.choice(P - B, 100, replace=False)])
    want = lmm.compat_rotate(K, X, Y[:, cols])
    ld_o = lmm.fit_delta_grid_compat(want, n_grid=n_grid)
    ldl = ld_l.cpu().numpy()[cols]
    loss = lambda v: lmm.profile_loss_compat_per_pheno(v, want["s"], want["UY"])  # noqa: E731
    gap = loss(ldl) - loss(ld_o)
    assert gap.max() <= 1e-6, gap.max()
    res = core.snp_wald_scan_compat(spec, ld_l, T(St), n_real=B)
    ref = lmm.snp_wald_scan_compat({**want, "UY": want["UY"][:, :100]}, ldl[:100], St)
    got = {k: getattr(res, k).cpu().numpy()[:, :100] for k in ("f", "beta", "se")}
    dlp, dbeta, drel = check_scan(got, ref, N - C - 1, what="learned delta N=2000")
    say(f"learned delta N={N}, {P} columns: fallback {fb.mean():.3f}, grid+Newton {t_n * 1e3:.1f} ms, learned "
        f"{t_l * 1e3:.1f} ms; max |d log delta| vs grid+Newton {d.max():.1e}; oracle loss gap {gap.max():.1e}; "
        f"scan |d log10 p| {dlp:.1e} ({time.perf_counter() - t0:.1f}s)")
