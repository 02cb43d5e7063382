"""GPU parity of the plain GLS path (rotate / fit_delta_grid / snp_wald_scan / loco_scan) against the FP64 oracle
and the reference's golden vectors."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import golden
from oracle import lmm
from test_gpu_compat import check_scan

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.fixture(scope="module")
def ex(example):
    return {**example, "Zs": lmm.standardise_columns(example["Z_raw"].astype(np.float64))}


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def test_loco_scan_vs_reference_golden(ex):
    from fasterlmm_amd import core
    ref = golden("plain_loco_cov2.npz")
    X = ex["X_cov2"]
    res = core.loco_scan(T(ex["Zs"]), T(X), T(ex["Y"][:, :3]), ex["chrom"])
    check_scan(res, ref, ex["Zs"].shape[0] - X.shape[1] - 1, what="plain loco")
    np.testing.assert_allclose(res.nullh2.cpu().numpy(), ref["nullh2"], rtol=1e-5)
    np.testing.assert_allclose(res.max_F.cpu().numpy(), ref["max_F"], rtol=1e-4)


@pytest.mark.parametrize("reml", [False, True])
def test_fit_delta_grid_ml_reml_vs_reference(ex, reml):
    from fasterlmm_amd import core
    ref = golden("plain_loco_cov2.npz")["log_delta_reml" if reml else "log_delta_ml"]
    Zs, chrom, X, Y = ex["Zs"], ex["chrom"], ex["X_cov2"], ex["Y"][:, :3]
    for ci, c in enumerate(sorted(np.unique(chrom))[:5]):
        K = lmm.grm(Zs[:, chrom != c])
        spec = core.rotate(T(K), T(X), T(Y))
        ld = core.fit_delta_grid(spec, reml=reml).cpu().numpy()
        assert np.abs(ld - ref[ci]).max() <= 1e-4, (ci, ld, ref[ci])
        g = core.fit_delta_grid(spec, reml=reml, refine=False).cpu().numpy()
        want = lmm.fit_delta_grid(lmm.rotate(K, X, Y), reml=reml, refine=False)
        np.testing.assert_allclose(g, want, atol=1e-12)


def test_snp_wald_scan_at_oracle_delta(ex):
    from fasterlmm_amd import core
    Zs, chrom, X, Y = ex["Zs"], ex["chrom"], ex["X_cov2"], ex["Y"][:, :4]
    K = lmm.grm(Zs[:, chrom != 2])
    want = lmm.rotate(K, X, Y)
    ld = lmm.fit_delta_grid(want)
    S = Zs[:, chrom == 2]
    ref = lmm.snp_wald_scan(want, ld, want["U"].T @ S, n_real=2)
    spec = core.rotate(T(K), T(X), T(Y))
    res = core.snp_wald_scan(spec, T(ld), spec.U.T @ T(S), n_real=2)
    check_scan(res, ref, Zs.shape[0] - X.shape[1] - 1, what="plain scan")
    np.testing.assert_allclose(res.max_F.cpu().numpy(), ref["max_F"], rtol=1e-4)


def test_single_k_scan_vs_oracle(ex):
    from fasterlmm_amd import core
    X = ex["X_cov2"]
    ref = lmm.single_k_scan(ex["Zs"], X, ex["Y"][:, :2])
    res = core.single_k_scan(T(ex["Zs"]), T(X), T(ex["Y"][:, :2]))
    check_scan(res, ref, ex["Zs"].shape[0] - X.shape[1] - 1, what="plain single-K")


def test_profile_loss_helpers_match_oracle(ex):
    from fasterlmm_amd import core
    Zs, X, Y = ex["Zs"], ex["X_cov2"], ex["Y"][:, :3]
    spec = lmm.rotate(lmm.grm(Zs), X, Y)
    grid = np.array([-4.0, -1.0, 0.0, 2.0, 5.0])
    for reml in (False, True):
        got = core._profile_loss(T(grid), T(spec["s"]), T(spec["X_rot"]), T(spec["Y_rot"]), reml=reml)
        np.testing.assert_allclose(got.cpu().numpy(), lmm.profile_loss(grid, spec["s"], spec["X_rot"],
                                                                       spec["Y_rot"], reml=reml), rtol=1e-10)
    cell = core.nLLeval(float(np.exp(0.5)), T(spec["s"]), T(spec["X_rot"]), T(spec["Y_rot"][:, 1]))
    assert abs(cell.item() - lmm.nll_eval(np.exp(0.5), spec["s"], spec["X_rot"], spec["Y_rot"][:, 1])) < 1e-8
