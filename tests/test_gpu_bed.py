"""GPU parity of the .bed decoder (flmm_bed_decode) and the reader API built on it, against the oracle."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import golden
from oracle import bed as obed, lmm

pytestmark = pytest.mark.gpu


def _write_trio(tmp_path, Z, iid, chrom, count_A1=True):
    prefix = tmp_path / "g"
    (tmp_path / "g.bed").write_bytes(obed.encode(Z, count_A1=count_A1).tobytes())
    (tmp_path / "g.fam").write_text("".join(f"F {s} 0 0 0 -9\n" for s in iid))
    (tmp_path / "g.bim").write_text("".join(f"{int(c)}\trs{j}\t0\t{1000 + j}\tA\tG\n" for j, c in enumerate(chrom)))
    return prefix


def _panel(n, m, seed, miss=0.05):
    rng = np.random.default_rng(seed)
    Z = rng.integers(0, 3, size=(n, m)).astype(np.float64)
    Z[rng.random((n, m)) < miss] = np.nan
    Z[:, 3] = 1.0          # constant column -> zeros
    Z[:, 5] = np.nan       # all missing -> zeros
    return Z


def test_decode_example_panel_all_output_kinds():
    from fasterlmm_amd import _lib
    from fasterlmm_amd.bed import _decode
    fx = golden("example_bed.npz")
    n, m = len(fx["fam"]), len(fx["bim"])
    want = obed.decode(fx["bed"], n, m)
    recs = fx["bed"][3:].reshape(m, (n + 3) // 4)
    dev = _lib.device()
    g = _decode(recs, n, None, _lib.BED_I8, True, dev).cpu().numpy()
    np.testing.assert_array_equal(g.astype(np.float64), want)
    z64, mean, std = _decode(recs, n, None, _lib.BED_F64, True, dev, stats=True)
    ref = lmm.standardise_columns(want)
    assert np.abs(z64.cpu().numpy() - ref).max() <= 1e-13
    np.testing.assert_allclose(mean.cpu().numpy(), np.nanmean(want, axis=0), rtol=1e-15)
    np.testing.assert_allclose(std.cpu().numpy(), np.nanstd(want, axis=0), rtol=1e-13)
    z32 = _decode(recs, n, None, _lib.BED_F32, True, dev).cpu().numpy()
    np.testing.assert_allclose(z32, ref, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("count_A1", [True, False])
def test_decode_rows_cols_missing(count_A1):
    from fasterlmm_amd import _lib
    from fasterlmm_amd.bed import _decode
    n, m = 37, 300
    Z = _panel(n, m, 3)
    raw = obed.encode(Z, count_A1=count_A1)
    recs = raw[3:].reshape(m, (n + 3) // 4)
    rng = np.random.default_rng(5)
    rows = np.sort(rng.choice(n, size=29, replace=False))
    cols = rng.choice(m, size=120, replace=False)
    dev = _lib.device()
    g = _decode(recs[cols], n, rows, _lib.BED_I8, count_A1, dev).cpu().numpy().astype(np.float64)
    g[g < 0] = np.nan
    np.testing.assert_array_equal(g, Z[np.ix_(rows, cols)])
    z = _decode(recs[cols], n, rows, _lib.BED_F64, count_A1, dev).cpu().numpy()
    ref = lmm.standardise_columns(Z[np.ix_(rows, cols)])
    assert np.abs(z - ref).max() <= 1e-13
    for j in np.where(np.isin(cols, [3, 5]))[0]:
        assert np.all(z[:, j] == 0.0)


def test_reader_api(tmp_path):
    from fasterlmm_amd import bed
    n, m = 41, 260
    Z = _panel(n, m, 11)
    iid = [f"s{i}" for i in range(n)]
    chrom = np.repeat(np.arange(1, 5), m // 4)
    prefix = _write_trio(tmp_path, Z, iid, chrom)
    g = bed.read_plink(prefix)
    np.testing.assert_array_equal(g.Z, Z)
    assert g.iid == iid and len(g.sid) == m
    np.testing.assert_array_equal(g.chrom, chrom.astype(np.float64))
    keep = [iid[i] for i in range(n) if i % 3 != 1][::-1]  # reversed: the bed's own order must win
    h = bed.open_aligned_bed(prefix, keep)
    rows = np.array([i for i in range(n) if i % 3 != 1])
    assert h.iid == [iid[i] for i in rows]
    cols = np.arange(17, 170)
    blk = bed.read_block(h, cols, dtype=torch.float64).cpu().numpy()
    ref = lmm.standardise_columns(Z[np.ix_(rows, cols)])
    assert np.abs(blk - ref).max() <= 1e-13
    b32 = bed.read_block(h, cols).cpu().numpy()
    assert b32.dtype == np.float32 and np.abs(b32 - ref).max() <= 1e-6
    var = bed.stream_genotype_var(h, block_size=64)
    with np.errstate(invalid="ignore"):
        want_var = np.nanvar(Z[rows], axis=0)  # NaN for the all-missing variant, as the reference's nanvar
    np.testing.assert_allclose(var, want_var, rtol=1e-12, atol=1e-15)
    allz = bed.read_all_standardised(h, block_size=64, dtype=torch.float64).cpu().numpy()
    assert np.abs(allz - lmm.standardise_columns(Z[rows])).max() <= 1e-13
    for blk2, cidx in bed.chrom_test_blocks(h, 2, block_size=20, dtype=torch.float64):
        assert np.all(chrom[cidx] == 2)
        assert np.abs(blk2.cpu().numpy() - lmm.standardise_columns(Z[np.ix_(rows, cidx)])).max() <= 1e-13
