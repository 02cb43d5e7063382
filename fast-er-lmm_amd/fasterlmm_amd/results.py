"""Association results in the reference's on-disk schema.

The per-phenotype table is the 14 fastlmm ``single_snp`` columns plus the permutation ``threshold`` and a
``significant`` flag, sorted by PValue (fasterlmm/cli.py:114-161): per-phenotype directories hold ``gwas.tsv``
(the 14 columns, bare unquoted TSV), ``perms.tsv`` and ``threshold.txt``; the bundle is a
directory of parquet parts with all 16 columns (fasterlmm/bundle.py:38-120).  The p-values and the threshold
come from the GPU F(1, df2) tail (``pvalues``); this module only lays the tables out.
"""
from __future__ import annotations

import shutil
import threading
from pathlib import Path

import numpy as np
import pyarrow as pa
import pyarrow.csv as pacsv
import pyarrow.parquet as pq

# fastlmm single_snp's column order (the reference's tests/conftest.py GWAS_SCHEMA_COLS)
GWAS_COLUMNS = ["sid_index", "SNP", "Chr", "GenDist", "ChrPos", "PValue", "SnpWeight", "SnpWeightSE",
                "EffectSize", "SnpFractVarExpl", "Mixing", "Nullh2", "Pheno", "PhenoCount"]
BUNDLE_PARTS_DIRNAME = ".bundle_parts"
BUNDLE_FILENAME = "gwas_bundle.parquet"
_TSV_WRITE_OPTS = pacsv.WriteOptions(delimiter="\t", include_header=False, quoting_style="none")


def write_tsv(table: pa.Table, path) -> None:
    """Bare unquoted TSV with a hand-written header line, the way fastlmm writes it (cli.py:103-112)."""
    with open(path, "wb") as fh:
        fh.write(("\t".join(table.column_names) + "\n").encode())
        pacsv.write_csv(table, fh, write_options=_TSV_WRITE_OPTS)


def writer_context(snp_id, chrom, pos, g_var, n_perm: int) -> dict:
    """The per-run constant columns (cli.py:247-252): SNP ids, Chr, ChrPos, raw genotype variance (NaN-aware,
    ddof 0) for EffectSize, sid_index, all-null GenDist, all-zero Mixing, PhenoCount = 1 + n_perm."""
    M = len(snp_id)
    return {"snp_id": pa.array(list(snp_id), type=pa.string()), "chrom": np.asarray(chrom, dtype=np.float64),
            "pos": np.asarray(pos, dtype=np.float64), "g_var": np.asarray(g_var, dtype=np.float64),
            "sid_index": np.arange(M, dtype=np.int64), "gendist": pa.nulls(M, pa.float64()),
            "mixing": np.zeros(M), "phenocount": np.full(M, 1 + int(n_perm)),
            "pheno_idx": pa.array(np.zeros(M, dtype=np.int32))}


def pheno_table(ctx: dict, pheno_name: str, p_col, beta_col, se_col, sfve_col, nullh2_col, p_var: float,
                perm_min_p, perm_quantile: float) -> tuple[pa.Table, float]:
    """One phenotype's 16-column table sorted by PValue and its threshold (cli.py:114-144)."""
    thresh = float(np.quantile(np.asarray(perm_min_p, dtype=np.float64), perm_quantile))
    beta_col = np.asarray(beta_col, dtype=np.float64)
    effect_size = beta_col * beta_col * ctx["g_var"] / p_var  # fastlmm single_snp.py:1454
    p_col = np.asarray(p_col, dtype=np.float64)
    cols = {"sid_index": ctx["sid_index"], "SNP": ctx["snp_id"], "Chr": ctx["chrom"], "GenDist": ctx["gendist"],
            "ChrPos": ctx["pos"], "PValue": p_col, "SnpWeight": beta_col,
            "SnpWeightSE": np.asarray(se_col, dtype=np.float64), "EffectSize": effect_size,
            "SnpFractVarExpl": np.asarray(sfve_col, dtype=np.float64), "Mixing": ctx["mixing"],
            "Nullh2": np.asarray(nullh2_col, dtype=np.float64),
            "Pheno": pa.DictionaryArray.from_arrays(ctx["pheno_idx"], [pheno_name]),
            "PhenoCount": ctx["phenocount"], "threshold": np.full(len(p_col), thresh), "significant": p_col < thresh}
    return pa.table(cols).sort_by([("PValue", "ascending")]), thresh


def write_pheno(ctx: dict, outdir, pheno_name: str, p_col, beta_col, se_col, sfve_col, nullh2_col, p_var: float,
                perm_min_p, perm_quantile: float, per_pheno_dirs: bool, bundle_writer) -> None:
    """Write one phenotype's results (cli.py:114-161): per-phenotype directory and/or one bundle row group."""
    table, thresh = pheno_table(ctx, pheno_name, p_col, beta_col, se_col, sfve_col, nullh2_col, p_var,
                                perm_min_p, perm_quantile)
    if per_pheno_dirs:
        sub = Path(outdir) / pheno_name
        sub.mkdir(parents=True, exist_ok=True)
        write_tsv(table.select(GWAS_COLUMNS), sub / "gwas.tsv")
        write_tsv(pa.table({"perm_min_p": np.asarray(perm_min_p, dtype=np.float64)}), sub / "perms.tsv")
        (sub / "threshold.txt").write_text(f"{thresh:.6e}\n")
    if bundle_writer is not None:
        bundle_writer.append(table)


def write_batch(ctx: dict, outdir, pheno_names, Y_batch, res, perm_max_F, df2: float, *,
                perm_quantile: float = 0.05, per_pheno_dirs: bool = True, bundle_writer=None) -> None:
    """Write a scanned batch (cli.py:400-425): p = F(1, df2).sf(F) for the real columns and the per-permutation
    minimum p from the genome-wide maximum F, both on the GPU, then one table per phenotype.  Y_batch is the
    (N, B) phenotype matrix the scan saw (its ddof-0 variance is EffectSize's denominator)."""
    from .pvalues import pvalue
    p_real = pvalue(res.f, df2).cpu().numpy()
    perm_min_p = pvalue(perm_max_F, df2).cpu().numpy()
    beta, se = res.beta.cpu().numpy(), res.se.cpu().numpy()
    sfve, nullh2 = res.sfve.cpu().numpy(), res.nullh2.cpu().numpy()
    Yb = Y_batch.cpu().numpy() if hasattr(Y_batch, "cpu") else np.asarray(Y_batch)
    p_var = Yb.var(axis=0)
    for j, name in enumerate(pheno_names):
        write_pheno(ctx, outdir, name, p_real[:, j], beta[:, j], se[:, j], sfve[:, j], nullh2[:, j], float(p_var[j]),
                    perm_min_p[j], perm_quantile, per_pheno_dirs, bundle_writer)


class BundleWriter:
    """A directory of parquet parts, one per writing thread, one row group per phenotype, staged in a ``.tmp``
    sibling and renamed into place on close (bundle.py:38-120).  Pheno is stored as plain strings."""

    def __init__(self, path) -> None:
        self._dir = Path(path)
        self._tmp = self._dir.with_name(self._dir.name + ".tmp")
        if self._tmp.is_dir():
            shutil.rmtree(self._tmp)
        elif self._tmp.exists():
            self._tmp.unlink()
        self._tmp.mkdir(parents=True)
        self._local = threading.local()
        self._lock = threading.Lock()
        self._writers: list[pq.ParquetWriter] = []
        self._next_part = 0

    def append(self, table: pa.Table) -> None:
        i = table.schema.get_field_index("Pheno")
        table = table.set_column(i, "Pheno", table.column("Pheno").cast(pa.string()))
        writer = getattr(self._local, "writer", None)
        if writer is None:
            with self._lock:
                writer = pq.ParquetWriter(self._tmp / f"gwas-results-part-{self._next_part:05d}.parquet",
                                          table.schema, compression="snappy")
                self._next_part += 1
                self._writers.append(writer)
            self._local.writer = writer
        writer.write_table(table)

    def close(self) -> Path:
        for w in self._writers:
            w.close()
        self._writers = []
        if self._dir.is_dir():
            shutil.rmtree(self._dir)
        elif self._dir.exists():
            self._dir.unlink()
        self._tmp.replace(self._dir)
        return self._dir
