"""GPU invariants the reference's own test suite pins for the scan (tests/test_core_compat.py, test_core_plain.py,
test_kinship_panel.py, test_perms.py), checked on the MI355X path.

* SNP blocking is invisible: any internal block size gives the same bits (test_core_compat.py:244-258).
* max_F of a real column is the column maximum of f; nullh2 = 1 / (1 + delta); the n_real split keeps the real
  columns' rows and only the maximum of the others (test_core_plain.py:153-202, test_core_compat.py:261-284).
* LOCO nullh2 is constant within a chromosome and varies between chromosomes (test_core_compat.py:306-334).
* duplicate phenotype columns get identical deltas; refine=False lands on a grid node (test_core_compat.py:148-186).
* a kinship panel equal to the genotypes reproduces the scan bit for bit, and poisoning chromosome c of the panel
  leaves chromosome c's results unchanged up to rounding (test_kinship_panel.py:29-65).
* permutations: same seed -> identical bits; batching phenotypes together or apart, batch order and the progress
  callback do not change results; seed_idx keys the shuffles on the full-table column (test_perms.py:26-182).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import lmm
from test_gpu_compat import check_scan

pytestmark = pytest.mark.gpu
DEV = "cuda"


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.fixture(scope="module")
def ex(example):
    return {**example, "Zs": lmm.standardise_columns(example["Z_raw"].astype(np.float64))}


def _same(a, b, what):
    for k in ("f", "beta", "se", "sfve", "nullh2", "max_F"):
        assert torch.equal(getattr(a, k), getattr(b, k)), f"{what}: {k}"


def test_snp_blocking_is_invisible(ex):
    from fasterlmm_amd import core, engine
    Zs, chrom, X, Y = ex["Zs"], ex["chrom"], ex["X_cov2"], ex["Y"]
    spec = core.fastlmm_compat_rotate(T(lmm.grm(Zs[:, chrom > 3])), T(X), T(Y))
    ld = core.fit_delta_grid_compat(spec)
    st = spec._gpu
    S = T(Zs[:, chrom <= 3])
    M = S.shape[1]
    ri = torch.arange(M, device=DEV)
    outs = []
    for block in (M, 128, 97, 300):
        out = engine.ScanOutputs.allocate(M, 7, st.P, DEV)
        st.scan(S, ri, 7, out, st.delta_for(ld), block=block)
        outs.append(out)
    for o in outs[1:]:
        _same(o, outs[0], "block size")
    # and the drop-in's chunk arguments are accepted and change nothing
    a = core.snp_wald_scan_compat(spec, ld, S, n_real=7)
    b = core.snp_wald_scan_compat(spec, ld, S, n_real=7, snp_chunk=64, pheno_chunk=3)
    _same(a, b, "snp_chunk / pheno_chunk")


def test_max_f_nullh2_and_n_real_split(ex):
    from fasterlmm_amd import core
    Zs, chrom, X, Y = ex["Zs"], ex["chrom"], ex["X_cov2"], ex["Y"]
    spec = core.fastlmm_compat_rotate(T(lmm.grm(Zs[:, chrom != 5])), T(X), T(Y))
    ld = core.fit_delta_grid_compat(spec)
    S = T(Zs[:, chrom == 5])
    full = core.snp_wald_scan_compat(spec, ld, S)
    split = core.snp_wald_scan_compat(spec, ld, S, n_real=6)
    P = Y.shape[1]
    assert full.f.shape == (S.shape[1], P) and split.f.shape == (S.shape[1], 6)
    # max_F is the column maximum of f (every column real), and the split keeps every column's maximum
    torch.testing.assert_close(full.max_F, full.f.max(dim=0).values, rtol=0, atol=0)
    torch.testing.assert_close(split.max_F, full.max_F, rtol=0, atol=0)
    for k in ("f", "beta", "se", "sfve", "nullh2"):
        assert torch.equal(getattr(split, k), getattr(full, k)[:, :6]), k
    torch.testing.assert_close(full.nullh2, (1.0 / (1.0 + ld.exp())).unsqueeze(0).expand_as(full.nullh2),
                               rtol=1e-15, atol=0)
    assert torch.all(full.f >= 0) and torch.isfinite(full.f).all()


def test_loco_nullh2_is_per_chromosome(ex):
    from fasterlmm_amd import core
    Zs, chrom, X, Y = ex["Zs"], ex["chrom"], ex["X_cov2"], ex["Y"][:, :3]
    res = core.loco_scan_compat(T(Zs), T(X), T(Y), chrom)
    h = res.nullh2.cpu().numpy()
    per = []
    for c in sorted(np.unique(chrom)):
        blk = h[chrom == c]
        assert np.all(blk == blk[0]), c
        per.append(blk[0])
    assert len({tuple(np.round(p, 12)) for p in per}) > 1


def test_duplicate_columns_and_grid_nodes(ex):
    from fasterlmm_amd import core, engine
    Zs, chrom, X, Y = ex["Zs"], ex["chrom"], ex["X_cov2"], ex["Y"]
    Yd = np.concatenate([Y[:, :5], Y[:, :5]], axis=1)
    spec = core.fastlmm_compat_rotate(T(lmm.grm(Zs[:, chrom != 1])), T(X), T(Yd))
    for mode in ("newton", "golden"):
        with engine.configure(delta=mode):
            ld = core.fit_delta_grid_compat(spec).cpu().numpy()
        np.testing.assert_array_equal(ld[:5], ld[5:])
        assert np.all((ld >= -10.0) & (ld <= 10.0))
    g = core.fit_delta_grid_compat(spec, refine=False, n_grid=64).cpu().numpy()
    grid = np.linspace(-10.0, 10.0, 64)
    assert np.all(np.min(np.abs(g[:, None] - grid[None, :]), axis=1) < 1e-12)


def test_kinship_panel_identity_and_poisoning(ex):
    from fasterlmm_amd import core
    Zs, chrom, X, Y = ex["Zs"], ex["chrom"], ex["X_cov2"], ex["Y"][:, :3]
    base = core.loco_scan_compat(T(Zs), T(X), T(Y), chrom)
    same = core.loco_scan_compat(T(Zs), T(X), T(Y), chrom, Z_kin=T(Zs), chrom_kin=chrom)
    _same(same, base, "kinship panel = genotypes")
    poisoned = Zs.copy()
    rng = np.random.default_rng(9)
    c = 2
    poisoned[:, chrom == c] = lmm.standardise_columns(rng.standard_normal((Zs.shape[0], int((chrom == c).sum()))))
    res = core.loco_scan_compat(T(Zs), T(X), T(Y), chrom, Z_kin=T(poisoned), chrom_kin=chrom)
    m = torch.from_numpy(chrom == c).to(DEV)
    # K_loco = (K_all - Z_c Z_c^T) / (M - M_c): the poisoned columns cancel exactly in arithmetic, up to the
    # rounding of K_all (the reference rebuilds K from the other chromosomes and is bitwise here).  The panel's
    # duplicate strains make K's eigenspaces degenerate, so that rounding rotates eigenvectors inside them: the
    # statistics are invariant to it up to the split operands' rounding, i.e. the scan's own bars.
    want = {k: getattr(base, k)[m].cpu().numpy() for k in ("f", "beta", "se")}
    check_scan(type(res)(*(getattr(res, k)[m] for k in ("f", "beta", "se", "sfve", "nullh2")), res.max_F), want,
               Zs.shape[0] - X.shape[1] - 1, what="poisoned kinship panel, its own chromosome")
    torch.testing.assert_close(res.nullh2[m], base.nullh2[m], rtol=1e-6, atol=0)
    assert float(((res.f[~m] - base.f[~m]).abs() / base.f[~m].clamp(min=1e-3)).max()) > 1e-2
    with pytest.raises(ValueError, match="go together"):
        core.loco_scan_compat(T(Zs), T(X), T(Y), chrom, Z_kin=T(Zs))


def test_permutation_determinism_batching_and_seed_idx(ex):
    from fasterlmm_amd.io import align_arrays
    from fasterlmm_amd.perms import perm_orders, perm_threshold
    n = ex["Z_raw"].shape[0]
    iid = [str(i) for i in range(n)]
    data = align_arrays(iid, ex["Z_raw"].astype(np.float64), iid, ex["Y"], iid, ex["X_cov2"][:, 1:],
                        chrom=ex["chrom"])
    r1, p1 = perm_threshold(data, [0, 1, 2], n_perm=3, seed=7)
    r2, p2 = perm_threshold(data, [0, 1, 2], n_perm=3, seed=7, on_chrom=lambda k, n: None)
    _same(r1, r2, "same seed, with callback")
    assert torch.equal(p1, p2)
    # batching: phenotype 1 alone, and the batch in another order
    ra, pa = perm_threshold(data, [1], n_perm=3, seed=7)
    torch.testing.assert_close(ra.f[:, 0], r1.f[:, 1], rtol=1e-7, atol=0)
    torch.testing.assert_close(pa[0], p1[1], rtol=1e-7, atol=0)
    rb, pb = perm_threshold(data, [2, 0, 1], n_perm=3, seed=7)
    torch.testing.assert_close(rb.f[:, [1, 2, 0]], r1.f, rtol=1e-7, atol=0)
    torch.testing.assert_close(pb[[1, 2, 0]], p1, rtol=1e-7, atol=0)
    # seed_idx keys the shuffles on the full-table column
    o1 = perm_orders(7, [5], n, 3)
    o2 = perm_orders(7, [6], n, 3)
    assert not np.array_equal(o1, o2)
    rc, pc = perm_threshold(data, [1], n_perm=3, seed=7, seed_idx=[1])
    assert torch.equal(pc, pa)
    rd, pd = perm_threshold(data, [1], n_perm=3, seed=7, seed_idx=[4])
    assert not torch.equal(pd, pa)
