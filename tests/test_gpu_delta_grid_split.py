"""The delta grid on split-FP16 MFMA (flmm_delta_grid_split) against the FP32 grid (flmm_delta_grid) and the oracle.

Both grids pick each column's minimum of the compat loss over the same 64 nodes (core.py:561-571); they may differ
only where two nodes' losses tie to their FP32 rounding (the loss, ~2e3 here, carries ~1e-4 of it), which is common:
a permuted column's loss flattens toward large delta (h2 -> 0), where every node of the tail ties.  The FP64 loss
of the oracle (profile_loss_compat) on the GPU's own spectrum and rotated columns judges every pick: each must lie
within 2e-3 + 2e-7 |loss| (in loss units; adjacent nodes differ by ~1 near a minimum) of the FP64 minimum.
The columns span 16 decades of scale (each carries its own plane scale) and the spectrum starts at 0 (the compat
clamp), where the row scales of the weights are largest.  End to end, a LOCO scan with either grid gives the same
delta fits and tests (to the refinement's tolerance) and matches the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import lmm, perms as operms

pytestmark = pytest.mark.gpu
DEV = "cuda"


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def panel():
    from fasterlmm_amd import sim
    cfg = dict(N=400, M=4000, n_chrom=2, P=4, C=3, maf=(0.05, 0.5), h2=(0.0, 0.9), rint=True)
    pan = sim.make_panel(cfg, seed=41)
    Zs = lmm.standardise_columns(pan.G.astype(np.float64))
    orders = operms.perm_orders(20240229, list(range(pan.Y.shape[1])), pan.Y.shape[0], 150)
    Y = operms.pack_perm_columns(pan.Y, orders)  # 604 columns
    return pan, Zs, Y


def _grids(st, n_grid=64, lo=-10.0, hi=10.0):
    from fasterlmm_amd import engine
    s_lam, nco = st._loss(True)
    args = (st.uy, st.P, st.p_pad, st.kpad, st.s, s_lam, nco, n_grid, lo, hi)
    split = engine.delta_grid(*args, col_scale=st.col_scale, s_min=st.s_range[0])[0].cpu().numpy()
    with engine.configure(grid="fp32"):
        fp32 = engine.delta_grid(*args, col_scale=st.col_scale, s_min=st.s_range[0])[0].cpu().numpy()
    return split, fp32


def test_split_grid_picks_the_fp64_minimum(panel):
    from fasterlmm_amd import core
    pan, Zs, Y = panel
    on = pan.chrom == 1
    K = lmm.grm(Zs[:, ~on])
    rng = np.random.default_rng(7)
    Ys = Y * 10.0 ** rng.uniform(-8, 8, size=Y.shape[1])  # every column at its own scale
    spec = core.fastlmm_compat_rotate(T(K), T(pan.X), T(Ys))
    st = spec._gpu
    assert st.col_scale is not None and st.s_range[0] >= 0.0
    split, fp32 = _grids(st)
    grid = np.linspace(-10.0, 10.0, 64)
    loss = lmm.profile_loss_compat(grid, st.s.cpu().numpy(), st.uy_dense().cpu().numpy())
    best = loss.min(axis=0)
    cols = np.arange(loss.shape[1])
    tol = 2e-3 + 2e-7 * np.abs(best)
    gap_s = (loss[split, cols] - best) / tol
    gap_f = (loss[fp32, cols] - best) / tol
    n_diff = int((split != fp32).sum())
    say(f"split grid: {n_diff} of {len(cols)} picks differ from the FP32 grid; FP64 loss gap max "
        f"{gap_s.max():.2e} (split), {gap_f.max():.2e} (FP32) of the tolerance")
    assert gap_s.max() <= 1.0, int(np.argmax(gap_s))
    assert gap_f.max() <= 1.0


def test_split_grid_pre_split_weights_same_picks(panel):
    """flmm_delta_grid_split with the grid rows' split planes formed once in a workspace (the engine's call) and
    formed in every column tile (no workspace): the same picks, bit for bit."""
    from fasterlmm_amd import _lib, core, engine
    pan, Zs, Y = panel
    on = pan.chrom == 1
    spec = core.fastlmm_compat_rotate(T(lmm.grm(Zs[:, ~on])), T(pan.X), T(Y))
    st = spec._gpu
    s_lam, nco = st._loss(True)
    idx_ws = engine.delta_grid(st.uy, st.P, st.p_pad, st.kpad, st.s, s_lam, nco, 64, -10.0, 10.0,
                               col_scale=st.col_scale, s_min=st.s_range[0])[0]
    wplane = torch.empty(128 * st.kpad, dtype=torch.float32, device=DEV)
    lam = torch.empty(128, dtype=torch.float32, device=DEV)
    lam64 = torch.empty(128, dtype=torch.float64, device=DEV)
    idx = torch.empty(st.P, dtype=torch.int32, device=DEV)
    stream = _lib.stream()
    _lib.call("flmm_delta_weights", st.s.data_ptr(), st.s.numel(), s_lam.data_ptr(), s_lam.numel(), st.kpad, 64,
              -10.0, 10.0, wplane.data_ptr(), lam.data_ptr(), lam64.data_ptr(), stream)
    _lib.call("flmm_delta_grid_split", wplane.data_ptr(), lam.data_ptr(), 64, float(nco), st.uy.data_ptr(),
              st.col_scale.data_ptr(), float(st.s_range[0]), -10.0, 10.0, st.P, st.p_pad, st.kpad, idx.data_ptr(),
              None, stream)
    torch.cuda.synchronize()
    assert torch.equal(idx, idx_ws)


def test_split_grid_loco_matches_fp32_grid_and_oracle(panel):
    from fasterlmm_amd import core, engine
    pan, Zs, Y = panel
    args = (T(Zs), T(pan.X), T(Y), pan.chrom)
    res = {}
    for mode in ("auto", "fp32"):
        with engine.configure(grid=mode):
            res[mode] = core.loco_scan_compat(*args, n_real=4)
    torch.cuda.synchronize()
    a, b = res["auto"], res["fp32"]
    dh = np.abs(a.nullh2.cpu().numpy() - b.nullh2.cpu().numpy()).max()
    df = np.abs(a.f.cpu().numpy() - b.f.cpu().numpy()) / np.maximum(np.abs(b.f.cpu().numpy()), 1e-300)
    mf = np.abs(a.max_F.cpu().numpy() - b.max_F.cpu().numpy()) / b.max_F.cpu().numpy()
    ref = lmm.loco_scan_compat(Zs, pan.X, Y, pan.chrom, n_real=4)
    dref = np.abs(a.max_F.cpu().numpy() - ref["max_F"]) / ref["max_F"]
    say(f"LOCO split vs FP32 grid: nullh2 {dh:.2e}, F rel {df.max():.2e}, max_F rel {mf.max():.2e}; "
        f"max_F vs oracle {dref.max():.2e}")
    assert dh <= 1e-6
    assert df.max() <= 1e-5
    assert mf.max() <= 1e-5
    assert dref.max() <= 1e-4
