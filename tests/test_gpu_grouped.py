"""GPU tests of the grouped scan: several column sets of uneven retained N in one persistent launch.

* the grouped launch (either tile variant, any grid size, eager or replayed from a captured HIP graph) writes
  exactly what one flmm_wald_scan launch per group writes;
* graph replays after the table is rewritten for other groups stay exact, with eager work between replays;
* the autotuner picks one of its candidates and its choice gives the same results;
* the NA-grouped permutation driver matches the FP64 oracle run per group on the group's strain subset
  (|d log10 p| <= 1e-3, |d beta| <= 1e-5 max(|beta|, se), perm max F to 1e-4).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import lmm, perms as operms
from test_gpu_compat import check_scan

pytestmark = pytest.mark.gpu

DEV = "cuda"
FIELDS = ("f", "beta", "se", "sfve", "nullh2", "max_F")


@pytest.fixture(scope="module")
def mods():
    from fasterlmm_amd import _lib, core, engine, grouped
    return _lib, core, engine, grouped


def _groups(example, n_groups=3, seed=4):
    """Column sets of uneven retained N from the example panel: group g keeps 150 - 25 g strains and scans
    20 + 90 g columns (reals plus shuffles); SNPs of chromosomes 1-4 (ragged SNP and column tiles)."""
    Zs_full = example["Z_raw"].astype(np.float64)
    chrom = example["chrom"]
    X = example["X_cov2"]
    Y = example["Y"]
    rng = np.random.default_rng(seed)
    out = []
    for g in range(n_groups):
        N = 150 - 25 * g
        keep = np.sort(rng.permutation(150)[:N])
        Zs = lmm.standardise_columns(Zs_full[keep])
        P = 20 + 90 * g
        Yg = np.concatenate([Y[keep]] + [Y[keep][rng.permutation(N)] for _ in range(P // 20)], axis=1)[:, :P]
        K = lmm.grm(Zs[:, chrom > 4])
        S = Zs[:, chrom <= 4][:, : 300 - 60 * g]
        out.append(dict(K=K, X=X[keep], Y=Yg, S=S, n_real=min(7 + g, P)))
    return out


def _jobs(mods, groups, precision):
    _lib, core, engine, grouped = mods
    jobs = []
    with engine.configure(precision=precision):
        for gr in groups:
            spec = core.fastlmm_compat_rotate(*(torch.from_numpy(gr[k]).to(DEV) for k in ("K", "X", "Y")))
            st = spec._gpu
            _, d = st.fit(64, -10.0, 10.0, True)
            S = torch.from_numpy(gr["S"]).to(DEV)
            mc = S.shape[1]
            ri = torch.arange(mc, device=DEV)
            s1, s2 = st.snp_planes(S, ri)
            out = core._new_outputs(mc, gr["n_real"], st.P, DEV)
            jobs.append(grouped.ScanJob(s1, s2, d, mc, st.P, gr["n_real"], st.nu, ri, out))
    return jobs


def _reset(jobs):
    for j in jobs:
        for name in FIELDS[:5]:
            getattr(j.out, name).fill_(float("nan"))
        j.out.max_F.fill_(float("-inf"))


def _snapshot(jobs):
    return [{k: getattr(j.out, k).clone() for k in FIELDS} for j in jobs]


def _assert_same(got, want, what):
    for g, (a, b) in enumerate(zip(got, want)):
        for k in FIELDS:
            assert torch.equal(a[k], b[k]), f"{what}: group {g} field {k} differs"


@pytest.mark.parametrize("precision", ["split_fp16", "split_bf16", "fp32"])
def test_grouped_equals_per_group_launches(mods, example, precision, monkeypatch):
    _lib, core, engine, grouped = mods
    jobs = _jobs(mods, _groups(example), precision)
    prec = engine.prec_code(precision)
    # the default weights: factored for the split precisions where the sum is shorter than Neff (the 150- and
    # 125-strain groups, K = 256), exact for the 100-strain group (K = 128) and for fp32: a mixed job list runs as
    # one launch per kind
    kinds = [j.factored for j in jobs]
    assert kinds == ([True, False, False] if precision != "fp32" else [False] * 3), kinds
    variants = ((_lib.SCAN_WIDE, None),) if any(kinds) else ((_lib.SCAN_WIDE, None), (_lib.SCAN_SQUARE, "square"))
    for variant, env in variants:
        if env:
            monkeypatch.setenv("FLMM_SCAN_VARIANT", env)
        else:
            monkeypatch.delenv("FLMM_SCAN_VARIANT", raising=False)
        _reset(jobs)
        grouped.per_group_loop(jobs)
        want = _snapshot(jobs)
        total = sum(j.tiles(variant) for j in jobs)
        for blocks in (8, 64, engine.rup(total, 8), 512):
            for graph in (False, True):
                runner = grouped.GroupedScan(prec, len(jobs), max(j.P for j in jobs), graph=graph, variant=variant,
                                             blocks=blocks)
                _reset(jobs)
                runner.run(jobs)
                torch.cuda.synchronize()
                _assert_same(_snapshot(jobs), want, f"{precision} variant {variant} blocks {blocks} graph {graph}")


def test_graph_replays_after_table_rewrites(mods, example):
    """One captured graph replays for different group lists (fewer groups, other sizes) with eager torch work and
    synchronisations between replays; every replay matches the per-group launches."""
    _lib, core, engine, grouped = mods
    jobs_a = _jobs(mods, _groups(example, 3, seed=4), "split_fp16")
    jobs_b = _jobs(mods, _groups(example, 2, seed=9)[::-1], "split_fp16")
    want = {}
    for name, jobs in (("a", jobs_a), ("b", jobs_b)):
        _reset(jobs)
        grouped.per_group_loop(jobs)
        want[name] = _snapshot(jobs)
    max_P = max(j.P for j in jobs_a + jobs_b)
    runner = grouped.GroupedScan(_lib.PREC_SPLIT_FP16, 4, max_P, graph=True, variant=_lib.SCAN_WIDE, blocks=256)
    x = torch.randn(512, 512, device=DEV, dtype=torch.float64)
    for rep, name in enumerate("ababba"):
        jobs = jobs_a if name == "a" else jobs_b
        _reset(jobs)
        runner.run(jobs)
        y = (x @ x).sum()  # eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(y)
        _assert_same(_snapshot(jobs), want[name], f"replay {rep} ({name})")
    # one graph per weight kind present (factored and exact jobs run as separate launches), reused by every replay
    kinds = {j.factored for j in jobs_a + jobs_b}
    assert len(runner._graphs) == len(kinds)


def test_autotuner_choice(mods, example, tmp_path):
    _lib, core, engine, grouped = mods
    jobs = _jobs(mods, _groups(example), "split_fp16")
    _reset(jobs)
    grouped.per_group_loop(jobs)
    want = _snapshot(jobs)
    path = str(tmp_path / "tune.json")
    tuner = grouped.ScanTuner(path=path)
    runner = grouped.GroupedScan(_lib.PREC_SPLIT_FP16, len(jobs), max(j.P for j in jobs), tuner=tuner)
    _reset(jobs)
    runner.run(jobs)
    torch.cuda.synchronize()
    parts = [p for p in ([j for j in jobs if j.factored], [j for j in jobs if not j.factored]) if p]
    assert len(runner.last_choices) == len(parts)
    got = _snapshot(jobs)
    tuner2 = grouped.ScanTuner(path=path)
    for part, choice in zip(parts, runner.last_choices):
        key = tuner.key(_lib.PREC_SPLIT_FP16, part)
        assert choice in tuner.candidates(part)
        assert len(tuner.trials[key]) == len(tuner.candidates(part))
        assert tuner2.cache[key] == choice  # the choice persists and is reused without new trials
    if all(c[0] in (_lib.SCAN_WIDE, _lib.SCAN_WIDE_FACTORED) for c in runner.last_choices):  # the loop's tile bodies
        _assert_same(got, want, "tuned")
    else:  # the square variant rounds the FP64 unscaling differently
        for a, b in zip(got, want):
            for k in FIELDS:
                torch.testing.assert_close(a[k], b[k], rtol=1e-12, atol=0)


def _na_dataset(example, plant: bool = False):
    """Four missingness patterns over 18 phenotypes, missing calls in Z; plant: two phenotypes of three patterns
    are chromosome-2 SNPs plus noise (F from ~60 to ~1.4e5, above the FP64 re-scoring threshold)."""
    from fasterlmm_amd.io import align_arrays
    n = example["Z_raw"].shape[0]
    Y = example["Y"].astype(np.float64).copy()
    rng = np.random.default_rng(21)
    miss1 = rng.permutation(n)[:30]
    miss2 = rng.permutation(n)[:55]
    Y[miss1, 6:12] = np.nan
    Y[miss2, 12:17] = np.nan
    Y[rng.permutation(n)[:10], 17] = np.nan
    iid = [str(i) for i in range(n)]
    Z = example["Z_raw"].astype(np.float64)
    Z[rng.integers(0, n, 60), rng.integers(0, Z.shape[1], 60)] = np.nan  # missing calls
    if plant:  # 13 tests above F = 60 on chromosome 2 (oracle), at several SNP positions
        Zs, on2 = lmm.standardise_columns(Z), np.flatnonzero(example["chrom"] == 2)
        noise = np.random.default_rng(8).standard_normal((n, 2))
        for j in (0, 6, 12):
            Y[:, j] = np.where(np.isnan(Y[:, j]), np.nan, Zs[:, on2[3]] + 0.005 * noise[:, 0])
            Y[:, j + 1] = np.where(np.isnan(Y[:, j + 1]), np.nan, Zs[:, on2[40]] + 0.3 * noise[:, 1])
    return align_arrays(iid, Z, iid, Y, iid, example["X_cov2"][:, 1:], chrom=example["chrom"]), Z, Y


@pytest.mark.parametrize("mode,loco", [("grouped", True), ("loop", True), ("grouped", False)])
def test_perm_threshold_groups_vs_oracle(example, mode, loco):
    from fasterlmm_amd.perms import perm_threshold_groups
    data, Z, Y = _na_dataset(example)
    pheno = list(range(18))
    res = perm_threshold_groups(data, pheno, n_perm=3, seed=19930909, mode=mode, loco=loco)
    groups = lmm.group_by_na(Y, pheno)
    assert [(k, c) for k, c, _, _ in res] == [(list(k), list(c)) for k, c in groups]
    assert len(res) == 4
    X = data.X.cpu().numpy()
    for keep, cols, r, pmf in res:
        ref, ref_pmf = operms.perm_threshold(Z[keep], X[keep], Y[keep][:, cols], example["chrom"],
                                             list(range(len(cols))), n_perm=3, seed=19930909, seed_idx=cols,
                                             loco=loco)
        df = len(keep) - X.shape[1] - 1
        check_scan(r, ref, df, what=f"group of {len(keep)} strains")
        np.testing.assert_allclose(pmf.cpu().numpy(), ref_pmf, rtol=1e-4)
        np.testing.assert_allclose(r.max_F.cpu().numpy(), ref["max_F"], rtol=1e-4)


@pytest.mark.parametrize("mode", ["grouped", "loop"])
def test_grouped_large_f_tests_are_rescored(example, mode):
    """Planted hits in three missingness patterns: every test above the re-scoring threshold equals the FP64
    oracle's Wald statistic of the group (its strains, its chromosome-2 LOCO kernel, its real columns) to 1e-8,
    evaluated at the driver's own log delta (recovered from nullh2: the driver's Newton refinement and the
    oracle's golden section land ~1e-7 apart where the likelihood is flat, which F follows); the rest matches the
    oracle's own run to the scan bars, and max_F of the hit columns follows their re-scored F."""
    from fasterlmm_amd.core import RESCORE_F
    from fasterlmm_amd.perms import perm_threshold_groups
    data, Z, Y = _na_dataset(example, plant=True)
    chrom = example["chrom"]
    on2 = chrom == 2
    res = perm_threshold_groups(data, list(range(18)), n_perm=3, seed=19930909, mode=mode)
    X = data.X.cpu().numpy()
    n_big = 0
    for keep, cols, r, pmf in res:
        what = f"group of {len(keep)} strains, planted hits"
        ref, ref_pmf = operms.perm_threshold(Z[keep], X[keep], Y[keep][:, cols], chrom, list(range(len(cols))),
                                             n_perm=3, seed=19930909, seed_idx=cols)
        check_scan(r, ref, len(keep) - X.shape[1] - 1, what=what)
        np.testing.assert_allclose(r.max_F.cpu().numpy(), ref["max_F"], rtol=1e-4)
        h2 = r.nullh2.cpu().numpy()[np.flatnonzero(on2)[0]]
        Zs = lmm.standardise_columns(Z[keep])
        spec = lmm.compat_rotate(lmm.grm(Zs[:, ~on2]), X[keep], Y[keep][:, cols])
        at = lmm.snp_wald_scan_compat(spec, np.log(1.0 / h2 - 1.0), Zs[:, on2])
        big = at["f"] > 2 * RESCORE_F
        n_big += int(big.sum())
        for k in ("f", "beta", "se"):
            np.testing.assert_allclose(getattr(r, k).cpu().numpy()[on2][big], at[k][big], rtol=1e-8,
                                       err_msg=f"{k}, {what}")
        hit = big.any(axis=0)
        np.testing.assert_allclose(r.max_F.cpu().numpy()[: len(cols)][hit], at["f"].max(axis=0)[hit], rtol=1e-8)
    assert n_big >= 12
