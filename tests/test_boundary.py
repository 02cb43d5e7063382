"""CPU tests of the drop-in boundary: the reference-named helpers its own tests import
(core._eigh_with_cpu_fallback, core._profile_loss_compat(_per_pheno), lowrank._profile_loss_lowrank(_per_pheno)),
io.align_inputs called with the reference's records as fasterlmm/cli.py:231 calls it, the table readers, and
the plain GLS path's refusal of a rank-deficient covariate design.  No GPU is used: the loss helpers are torch
FP64 on the inputs' device, and the rank check runs before any device work."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import lmm
from oracle import lowrank as olr


def _spd(n: int, seed: int = 19930909) -> torch.Tensor:
    A = np.random.default_rng(seed).normal(size=(n, n + 5))
    return torch.from_numpy(A @ A.T / (n + 5))


# ---- core._eigh_with_cpu_fallback (core.py:54-77; the reference's tests/test_core_fallback.py cases) ----

def test_eigh_fallback_reference_name_oom_then_retry(monkeypatch):
    from fasterlmm_amd.core import _eigh_with_cpu_fallback
    K = _spd(20)
    real = torch.linalg.eigh
    calls = {"n": 0}

    def flaky(A):
        calls["n"] += 1
        if calls["n"] == 1:
            raise RuntimeError("HIP out of memory. Tried to allocate ...")
        return real(A)

    monkeypatch.setattr(torch.linalg, "eigh", flaky)
    s, U = _eigh_with_cpu_fallback(K)
    assert calls["n"] == 2
    assert torch.allclose(U @ torch.diag(s) @ U.T, K, atol=1e-9)
    assert torch.all(s[:-1] <= s[1:])


def test_eigh_fallback_reference_name_other_errors_and_clean_path(monkeypatch):
    from fasterlmm_amd.core import _eigh_with_cpu_fallback
    real = torch.linalg.eigh
    calls = {"n": 0}

    def counted(A):
        calls["n"] += 1
        return real(A)

    monkeypatch.setattr(torch.linalg, "eigh", counted)
    K = _spd(15)
    s, U = _eigh_with_cpu_fallback(K)
    assert calls["n"] == 1 and torch.allclose(U @ torch.diag(s) @ U.T, K, atol=1e-9)

    def boom(A):
        raise RuntimeError("some unrelated linalg failure")

    monkeypatch.setattr(torch.linalg, "eigh", boom)
    with pytest.raises(RuntimeError, match="unrelated"):
        _eigh_with_cpu_fallback(_spd(12))


# ---- loss helpers against the oracle (itself pinned by dense slogdet in test_oracle*.py) ----

@pytest.fixture(scope="module")
def compat_spec(example):
    Zs = lmm.standardise_columns(example["Z_raw"].astype(np.float64))
    K = lmm.grm(Zs[:, example["chrom"] != 3])
    return lmm.compat_rotate(K, example["X_cov2"], example["Y"][:, :5])


def test_profile_loss_compat_matches_oracle(compat_spec):
    from fasterlmm_amd.core import _profile_loss_compat, _profile_loss_compat_per_pheno
    s, UY = compat_spec["s"], compat_spec["UY"]
    grid = np.linspace(-10, 10, 17)
    got = _profile_loss_compat(torch.from_numpy(grid), torch.from_numpy(s), torch.from_numpy(UY)).numpy()
    np.testing.assert_allclose(got, lmm.profile_loss_compat(grid, s, UY), rtol=1e-13)
    ld = np.array([-3.0, 0.5, 2.0, 7.0, -9.0])
    got = _profile_loss_compat_per_pheno(torch.from_numpy(ld), torch.from_numpy(s), torch.from_numpy(UY)).numpy()
    np.testing.assert_allclose(got, lmm.profile_loss_compat_per_pheno(ld, s, UY), rtol=1e-13)
    # per-pheno column p at log delta ld[p] equals the grid form's cell
    for p in range(5):
        cell = _profile_loss_compat(torch.tensor([ld[p]]), torch.from_numpy(s), torch.from_numpy(UY[:, p:p + 1]))
        assert abs(float(cell[0, 0]) - got[p]) <= 1e-10 * abs(got[p])


def test_profile_loss_lowrank_matches_oracle():
    from fasterlmm_amd.lowrank import _profile_loss_lowrank, _profile_loss_lowrank_per_pheno
    rng = np.random.default_rng(4)
    k, P, Neff = 12, 6, 40
    s = np.sort(rng.uniform(0.01, 5.0, k))
    UY = rng.normal(size=(k, P))
    yy_off = rng.uniform(1.0, 30.0, P)
    grid = np.linspace(-6, 6, 9)
    t = torch.from_numpy
    got = _profile_loss_lowrank(t(grid), t(s), t(UY), t(yy_off), Neff).numpy()
    np.testing.assert_allclose(got, olr.profile_loss_lowrank(grid, s, UY, yy_off, Neff), rtol=1e-13)
    ld = rng.uniform(-5, 5, P)
    got = _profile_loss_lowrank_per_pheno(t(ld), t(s), t(UY), t(yy_off), Neff).numpy()
    np.testing.assert_allclose(got, olr.profile_loss_lowrank_per_pheno(ld, s, UY, yy_off, Neff), rtol=1e-13)


# ---- io.align_inputs with the reference's records (io.py:143-184; tests/test_io.py:174-250 cases) ----

def test_align_inputs_intersection_reorder_and_dtype():
    from fasterlmm_amd.io import Genotypes, Phenotypes, Covariates, align_inputs
    Z = np.array([[0.0, 1.0], [1.0, 2.0], [2.0, 0.0]])  # strains A B C
    g = Genotypes(Z=Z, iid=["A", "B", "C"], sid=["s0", "s1"], chrom=np.array([1, 1]), pos=np.array([10, 20]))
    p = Phenotypes(iid=["C", "A", "D"], names=["t0"], Y=np.array([[30.0], [10.0], [99.0]]))
    ds = align_inputs(g, p)
    assert ds.iid == ["A", "C"]
    assert torch.equal(ds.Y.squeeze(1), torch.tensor([10.0, 30.0], dtype=torch.float64))
    assert torch.equal(ds.Z, torch.tensor([[0.0, 1.0], [2.0, 0.0]], dtype=torch.float64))
    assert ds.X.shape == (2, 1) and torch.all(ds.X == 1)
    assert ds.snp_id == ["s0", "s1"] and ds.pheno_names == ["t0"] and list(ds.pos) == [10, 20]
    c = Covariates(iid=["C", "B", "A"], names=["cov_0"], C=np.array([[3.0], [2.0], [1.0]]))
    ds = align_inputs(g, p, c, dtype=torch.float32)
    assert ds.X.dtype == ds.Y.dtype == ds.Z.dtype == torch.float32
    assert torch.equal(ds.X, torch.tensor([[1.0, 1.0], [1.0, 3.0]]))
    with pytest.raises(ValueError):
        align_inputs(g, Phenotypes(iid=["X"], names=["t0"], Y=np.zeros((1, 1))))


def test_align_inputs_kinship_panel():
    from fasterlmm_amd.io import Genotypes, Phenotypes, align_inputs
    rng = np.random.default_rng(0)
    iid = [f"s{i}" for i in range(6)]
    Zt, Zk = rng.integers(0, 3, (6, 5)).astype(float), rng.integers(0, 3, (6, 7)).astype(float)
    perm = [3, 0, 5, 1, 4, 2]
    geno = Genotypes(Z=Zt, iid=iid, sid=list("abcde"), chrom=np.ones(5), pos=np.arange(5))
    kin = Genotypes(Z=Zk[perm], iid=[iid[i] for i in perm], sid=list("vwxyzuq"), chrom=np.ones(7), pos=np.arange(7))
    pheno = Phenotypes(iid=iid, names=["p"], Y=rng.normal(size=(6, 1)))
    al = align_inputs(geno, pheno, None, kin=kin)
    assert al.iid == iid and np.array_equal(al.Z_kin.numpy(), Zk) and al.n_dropped_kin == 0
    kin4 = Genotypes(Z=Zk[:4], iid=iid[:4], sid=list("xyz0123")[:7], chrom=np.ones(7), pos=np.arange(7))
    al = align_inputs(geno, pheno, None, kin=kin4)
    assert al.iid == iid[:4] and al.n_dropped_kin == 2


def test_read_phen_and_covar(tmp_path):
    from fasterlmm_amd.io import read_covar, read_phen
    (tmp_path / "ph.tsv").write_text("Strain\tA\tB\ns1\t1.5\tunknown\ns2\t\t2.0\ns3\t-0.5\t 3.0\n")
    ph = read_phen(tmp_path / "ph.tsv")
    assert ph.iid == ["s1", "s2", "s3"] and ph.names == ["A", "B"]
    np.testing.assert_array_equal(np.isnan(ph.Y), [[False, True], [True, False], [False, False]])
    assert ph.Y[2, 1] == 3.0 and ph.Y[0, 0] == 1.5
    (tmp_path / "bad.tsv").write_text("Strain\tA\ns1\tyes\ns2\tno\n")
    with pytest.raises(ValueError, match="text"):
        read_phen(tmp_path / "bad.tsv")
    (tmp_path / "cov.tab").write_text("f1 s1 0.5 1\nf2 s2 -1 0\n")
    cv = read_covar(tmp_path / "cov.tab")
    assert cv.iid == ["s1", "s2"] and cv.names == ["cov_0", "cov_1"]
    np.testing.assert_array_equal(cv.C, [[0.5, 1.0], [-1.0, 0.0]])


# ---- the plain GLS path refuses a rank-deficient covariate design (core.py:166-171, 336-340) ----

def test_plain_path_rejects_rank_deficient_X(example):
    from fasterlmm_amd import core
    X_full = example["X_full"]  # intercept + 43 one-hot covariates: rank 43 of 44 columns
    assert np.linalg.matrix_rank(X_full) == X_full.shape[1] - 1
    Z = torch.zeros(X_full.shape[0], 4, dtype=torch.float64)
    Y = torch.zeros(X_full.shape[0], 2, dtype=torch.float64)
    with pytest.raises(ValueError, match="full-column-rank"):
        core.loco_scan(Z, torch.from_numpy(X_full), Y, np.array([1, 1, 2, 2]))
    with pytest.raises(ValueError, match="full-column-rank"):
        core.single_k_scan(Z, torch.from_numpy(X_full), Y)
    with pytest.raises(ValueError, match="full-column-rank"):
        core.rotate(torch.eye(X_full.shape[0], dtype=torch.float64), torch.from_numpy(X_full), Y)
    # the oracle shows why: the reference's Cholesky of X~^T V^-1 X~ fails at some grid deltas and not at others
    Zs = lmm.standardise_columns(example["Z_raw"].astype(np.float64))
    spec = lmm.rotate(lmm.grm(Zs[:, example["chrom"] != 1]), X_full, example["Y"][:, :2])
    loss = lmm.profile_loss(np.linspace(-10, 10, 64), spec["s"], spec["X_rot"], spec["Y_rot"])
    bad = np.isinf(loss[:, 0])
    assert bad.any() and not bad.all()
