"""The certified permutation-column screen (flmm_wald_scan_screened) against the FP64 oracle and the full split scan.

The screen replaces the scan of the column tiles past the real columns: their only output is max_F, which must be the
maximum of FP64 evaluations of the tests that can hold it.  At a fixed delta (the oracle's), max_F is compared with
the oracle's to 1e-5 relative (the perm columns' rotation leaves ~2e-6), on SNP sets built so that every column's
maximum has near-ties (twin SNPs whose F differs by ~1e-4 - 1e-2 relative, inside the screen's error band of ~1%):
a wrong pick among them would show.  End to end
(LOCO, the GPU's own delta), max_F matches the unscreened scan and the oracle, the real columns' outputs are
unchanged bit for bit, and several SNP blocks per chromosome (the running maximum) give the same result.
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
    cfg = dict(N=400, M=6000, n_chrom=2, P=4, C=3, maf=(0.05, 0.5), h2=(0.0, 0.9), rint=True)
    pan = sim.make_panel(cfg, seed=23)
    Zs = lmm.standardise_columns(pan.G.astype(np.float64))
    n_perm = 300
    orders = operms.perm_orders(19930909, list(range(pan.Y.shape[1])), pan.Y.shape[0], n_perm)
    Y = operms.pack_perm_columns(pan.Y, orders)  # 4 reals, 1,200 permutations: columns 256.. are screened
    return pan, Zs, Y


def _twins(S, rng):
    """A near-duplicate of every SNP, s + eps |s| z / sqrt(N) with z ~ N(0, 1) and eps log-uniform in [1e-4, 1e-2]:
    each column's maximum gets near-ties."""
    N, M = S.shape
    eps = 10.0 ** rng.uniform(-4, -2, size=M)
    return S + eps * np.linalg.norm(S, axis=0) * rng.standard_normal(S.shape) / np.sqrt(N)


@pytest.mark.parametrize("layout", ["apart", "interleaved"])
@pytest.mark.parametrize("weights", ["auto", "exact"])
def test_screen_max_f_at_fixed_delta(panel, layout, weights):
    from fasterlmm_amd import core, engine
    pan, Zs, Y = panel
    on = pan.chrom == 1
    K = lmm.grm(Zs[:, ~on])
    S0 = Zs[:, on][:, :1500]
    rng = np.random.default_rng(5)
    tw = _twins(S0, rng)
    if layout == "apart":  # twins 1,500 columns (~12 tiles) away
        S = np.concatenate([S0, tw], axis=1)
    else:                  # twins side by side, mostly in one tile
        S = np.empty((S0.shape[0], 2 * S0.shape[1]))
        S[:, 0::2], S[:, 1::2] = S0, tw
    want = lmm.compat_rotate(K, pan.X, Y)
    ld = lmm.fit_delta_grid_compat(want, n_grid=64)
    ref = lmm.snp_wald_scan_compat(want, ld, S, n_real=4)
    engine.screen_stats.reset(True)
    with engine.configure(weights=weights):
        spec = core.fastlmm_compat_rotate(T(K), T(pan.X), T(Y))
        res = core.snp_wald_scan_compat(spec, T(ld), T(S), n_real=4)
        mf = res.max_F.cpu().numpy()
    st = engine.screen_stats.summary()
    engine.screen_stats.reset(False)
    rel = np.abs(mf - ref["max_F"]) / ref["max_F"]
    say(f"screen at fixed delta ({layout}, weights={weights}): max_F rel err {rel[256:].max():.2e} screened, "
        f"{rel[:256].max():.2e} full; {st['fp64_evaluations']} FP64 evaluations of {st['screened_tests']} tests")
    assert st["screened_tests"] == S.shape[1] * (Y.shape[1] - 256)
    assert rel[256:].max() <= 1e-5, (rel[256:].max(), 256 + int(np.argmax(rel[256:])))
    assert rel[:256].max() <= 1e-5


@pytest.mark.parametrize("layout", ["spread", "edges", "clustered"])
def test_bucketed_screen_at_fixed_delta(panel, layout):
    """The bucketed screen (every screened column on the sWy-only tile, sWs interpolated within its log-delta
    bucket from FP64 node sums) against the general tile (both products on hi planes) and the oracle, with the
    screened columns' log delta spread over [-10, 10] (100 buckets of 1/5), placed on the buckets' nodes and edges
    (u = 0, 1/2, 1 and just inside), or half at exactly 10 (the grid's upper end, where null columns settle).
    max_F equals the general tile's bit for bit (both are the FP64 value of the same maximising test) and the
    oracle's."""
    from fasterlmm_amd import core, engine
    pan, Zs, Y = panel
    on = pan.chrom == 1
    K = lmm.grm(Zs[:, ~on])
    S0 = Zs[:, on][:, :1500]
    S = np.concatenate([S0, _twins(S0, np.random.default_rng(9))], axis=1)
    if layout == "spread":  # 2,900 SNPs: 23 tiles of 128 rows, so a 256-row workgroup holds only 128 rows
        S = S[:, :2900]
    want = lmm.compat_rotate(K, pan.X, Y)
    ld = lmm.fit_delta_grid_compat(want, n_grid=64).copy()
    cols = np.arange(256, Y.shape[1])
    rng = np.random.default_rng(17)
    if layout == "spread":
        ld[cols] = rng.uniform(-10.0, 10.0, cols.size)
    elif layout == "edges":
        b = rng.integers(0, 100, cols.size)
        u = rng.choice([0.0, 0.5, 1.0, 1e-12, 1.0 - 1e-12], cols.size)
        ld[cols] = -10.0 + (b + u) * 0.2
    else:
        ld[cols[: cols.size // 2]] = 10.0
    ld[256], ld[257] = -10.0, 10.0  # the range [-10, 10] whatever the draw
    ref = lmm.snp_wald_scan_compat(want, ld, S, n_real=4)
    got, ev, st = {}, {}, {}
    for mode in ("auto", "general"):
        engine.screen_stats.reset(True)
        with engine.configure(screen=mode):
            spec = core.fastlmm_compat_rotate(T(K), T(pan.X), T(Y))
            got[mode] = core.snp_wald_scan_compat(spec, T(ld), T(S), n_real=4).max_F.cpu().numpy()
        st[mode] = engine.screen_stats.summary()
        engine.screen_stats.reset(False)
    rel = np.abs(got["auto"] - ref["max_F"]) / ref["max_F"]
    say(f"bucketed screen ({layout}): max_F rel err {rel[256:].max():.2e}; FP64 evaluations "
        f"{st['auto']['fp64_evaluations']} (bucketed) vs {st['general']['fp64_evaluations']} (general tile)")
    assert st["auto"]["swy_only_tests"] == st["auto"]["screened_tests"] == S.shape[1] * cols.size
    assert st["general"]["swy_only_tests"] == 0
    np.testing.assert_array_equal(got["auto"], got["general"])
    assert rel[256:].max() <= 1e-5, (rel[256:].max(), 256 + int(np.argmax(rel[256:])))


def test_bucketed_screen_forces_columns_outside_its_buckets(panel):
    """Columns whose log delta lies outside the bucketed layout's range (here a layout over [-10, -2] for columns
    spread over [-10, 10]) are screened as forced: every row of theirs is re-evaluated in FP64, so max_F still equals
    the general tile's bit for bit, with more FP64 evaluations."""
    from fasterlmm_amd import _lib, core, engine
    pan, Zs, Y = panel
    on = pan.chrom == 1
    K = lmm.grm(Zs[:, ~on])
    S = Zs[:, on][:, :1000]
    want = lmm.compat_rotate(K, pan.X, Y)
    ld = lmm.fit_delta_grid_compat(want, n_grid=64).copy()
    cols = np.arange(256, Y.shape[1])
    ld[cols] = np.random.default_rng(3).uniform(-10.0, 10.0, cols.size)
    spec = core.fastlmm_compat_rotate(T(K), T(pan.X), T(Y))
    st = spec._gpu
    hf = engine.screen_start(st.prec, 4, st.P)
    fac = st.factored(float(ld.min()), float(ld.max()))
    got, ev = {}, {}
    for mode, t_range in (("general", None), ("auto", (-10.0, -2.0))):
        with engine.configure(screen=mode):
            d = engine.delta_refine(st.prec, _lib.DELTA_FIXED, st.uy, st.P, st.p_pad, st.kpad, st.s, st.s,
                                    float(st.neff), 1, 0.0, 0.0, None, log_delta_in=T(ld), fac=fac, hi_from=hf,
                                    t_range=t_range)
            assert (d.layout is not None) == (mode == "auto")
            out = engine.ScanOutputs.allocate(S.shape[1], 4, st.P, T(S).device)
            engine.screen_stats.reset(True)
            st.scan(T(S), torch.arange(S.shape[1], device=DEV), 4, out, d)
            engine.refresh_real_max(out, 4)
            got[mode] = out.max_F.cpu().numpy()
            ev[mode] = engine.screen_stats.summary()["fp64_evaluations"]
            engine.screen_stats.reset(False)
    say(f"forced columns: FP64 evaluations {ev['auto']} (layout over [-10, -2]) vs {ev['general']} (general tile)")
    np.testing.assert_array_equal(got["auto"], got["general"])
    assert ev["auto"] > ev["general"]


def test_screen_loco_matches_full_scan_and_oracle(panel):
    from fasterlmm_amd import core, engine
    pan, Zs, Y = panel
    args = (T(Zs), T(pan.X), T(Y), pan.chrom)
    res = {}
    for mode in ("auto", "general", "off"):
        with engine.configure(screen=mode):
            res[mode] = core.loco_scan_compat(*args, n_real=4)
    torch.cuda.synchronize()
    # the bucketed screen changes which tests are re-evaluated, not the maximum
    np.testing.assert_array_equal(res["auto"].max_F.cpu().numpy(), res["general"].max_F.cpu().numpy())
    # the real columns run the full scan either way: same bits
    for k in ("f", "beta", "se", "sfve", "nullh2"):
        np.testing.assert_array_equal(getattr(res["auto"], k).cpu().numpy(), getattr(res["off"], k).cpu().numpy())
    a, b = res["auto"].max_F.cpu().numpy(), res["off"].max_F.cpu().numpy()
    np.testing.assert_array_equal(a[:256], b[:256])
    rel_full = np.abs(a - b) / b
    ref = lmm.loco_scan_compat(Zs, pan.X, Y, pan.chrom, n_real=4)
    rel = np.abs(a - ref["max_F"]) / ref["max_F"]
    say(f"screened LOCO: max_F vs full scan {rel_full.max():.2e}, vs oracle {rel.max():.2e}")
    assert rel_full.max() <= 1e-5
    assert rel.max() <= 1e-4


def test_screen_running_maximum_over_blocks(panel):
    """Several SNP blocks per chromosome: each block's resolution starts from the running max_F of the earlier
    ones; the result equals one block's."""
    from fasterlmm_amd import core, engine
    pan, Zs, Y = panel
    on = pan.chrom == 2
    K = lmm.grm(Zs[:, ~on])
    want = lmm.compat_rotate(K, pan.X, Y)
    ld = T(lmm.fit_delta_grid_compat(want, n_grid=64))
    spec = core.fastlmm_compat_rotate(T(K), T(pan.X), T(Y))
    st = core._spec_state(spec)
    S = T(Zs[:, on])
    ri = torch.arange(S.shape[1], device=DEV, dtype=torch.int64)
    got = []
    for block in (8192, 640, 256):
        out = engine.ScanOutputs.allocate(S.shape[1], 4, Y.shape[1], DEV)
        st.scan(S, ri, 4, out, st.delta_for(ld), block=block)
        got.append(out.max_F.cpu().numpy())
    for g in got[1:]:
        np.testing.assert_allclose(g, got[0], rtol=1e-12, atol=0)
