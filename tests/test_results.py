"""CPU tests of the result writer: the reference's column schema, sort order, threshold, EffectSize, TSV layout
and the parquet bundle (fasterlmm/cli.py:103-161, fasterlmm/bundle.py:38-120)."""
from __future__ import annotations

import numpy as np
import pyarrow as pa
import pyarrow.parquet as pq

from fasterlmm_amd import results


def _ctx_and_cols(M=40, seed=0):
    rng = np.random.default_rng(seed)
    ctx = results.writer_context([f"rs{i}" for i in range(M)], np.repeat([1.0, 2.0], M // 2),
                                 np.arange(M) * 10.0, rng.uniform(0.1, 0.5, M), n_perm=9)
    cols = dict(p_col=rng.uniform(0, 1, M), beta_col=rng.standard_normal(M), se_col=rng.uniform(0.1, 1, M),
                sfve_col=rng.uniform(0, 1, M), nullh2_col=np.full(M, 0.3), p_var=1.7,
                perm_min_p=rng.uniform(0, 0.2, 9), perm_quantile=0.05)
    return ctx, cols


def test_pheno_table_schema_sort_threshold_effect_size():
    ctx, c = _ctx_and_cols()
    t, thr = results.pheno_table(ctx, "YAL001C", **c)
    assert t.column_names == results.GWAS_COLUMNS + ["threshold", "significant"]
    p = t.column("PValue").to_numpy()
    assert np.all(np.diff(p) >= 0)
    assert thr == np.quantile(c["perm_min_p"], 0.05)
    order = np.argsort(c["p_col"], kind="stable")
    np.testing.assert_array_equal(t.column("sid_index").to_numpy(), order)
    np.testing.assert_allclose(t.column("EffectSize").to_numpy(),
                               (c["beta_col"] ** 2 * ctx["g_var"] / c["p_var"])[order], rtol=1e-15)
    np.testing.assert_array_equal(t.column("significant").to_numpy(), p < thr)
    assert t.column("GenDist").null_count == len(p)
    assert set(t.column("PhenoCount").to_numpy().tolist()) == {10}


def test_per_pheno_dirs_and_bundle(tmp_path):
    ctx, c = _ctx_and_cols()
    bw = results.BundleWriter(tmp_path / results.BUNDLE_FILENAME)
    for name in ("A", "B"):
        results.write_pheno(ctx, tmp_path, name, **c, per_pheno_dirs=True, bundle_writer=bw)
    out = bw.close()
    lines = (tmp_path / "A" / "gwas.tsv").read_text().splitlines()
    assert lines[0].split("\t") == results.GWAS_COLUMNS
    assert len(lines) == 41 and '"' not in "".join(lines)
    assert (tmp_path / "A" / "threshold.txt").read_text() == f"{np.quantile(c['perm_min_p'], 0.05):.6e}\n"
    assert (tmp_path / "A" / "perms.tsv").read_text().splitlines()[0] == "perm_min_p"
    tab = pq.read_table(out)
    assert tab.column_names == results.GWAS_COLUMNS + ["threshold", "significant"]
    assert tab.num_rows == 80 and tab.column("Pheno").type == pa.string()
    assert sorted(set(tab.column("Pheno").to_pylist())) == ["A", "B"]
