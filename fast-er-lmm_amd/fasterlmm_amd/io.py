"""Drop-in scan-input helpers of the reference's fasterlmm/io.py on MI355X: the input records (Genotypes,
Phenotypes, Covariates) and their readers, the aligned dataset, column standardisation (a gfx950 kernel), the
kinship Gram (the FP64 Gram kernel), NA-pattern grouping and strain subsets.

``align_inputs(geno, pheno, covar=None, *, dtype, kin)`` takes the reference's records as its callers pass them
(fasterlmm/cli.py:231, the reference's tests); ``align_arrays`` is the same intersection for callers that hold
plain arrays.  PLINK .bed reading (``read_plink``) decodes the 2-bit records on the GPU (bed.py).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch
from torch import Tensor

from . import _lib
from .bed import Genotypes, read_plink  # noqa: F401  (io.py:24-35, 54-67 of the reference)


@dataclass
class Phenotypes:
    """Wide phenotype table (io.py:38-51): iid = N strains, names = P labels, Y (N, P) float64 with NaN for
    missing; codebook / auto_categorical describe text columns coded to numbers."""
    iid: list
    names: list
    Y: np.ndarray
    codebook: dict = field(default_factory=dict)
    auto_categorical: list = field(default_factory=list)


@dataclass
class Covariates:
    """PLINK-style covariate table (io.py:97-106): C (N, K) with columns named cov_0, cov_1, ..."""
    iid: list
    names: list
    C: np.ndarray


# extra missing-value tokens of a phenotype table, on top of pandas' defaults (normalize.py:74)
_TEXT_NA = ["unknown", "Unknown", "UNKNOWN", "?", "n/a", "N/A", "none given", "-"]


def read_phen(path, levels: dict | None = None) -> Phenotypes:
    """Wide phenotype TSV with a 'Strain' first column (io.py:70-94).  Numeric columns become Y (a column that
    pandas read as text only because of stray blanks is converted, as in the reference).  Text-coded phenotype
    columns (the reference's normalize.encode_text_column: yes/no tiers, --levels, one-vs-rest expansion) are
    phenotype preparation outside the scan and raise ValueError here; code them numerically first."""
    import pandas as pd

    df = pd.read_csv(path, sep="\t", na_values=_TEXT_NA)
    iid = df["Strain"].astype(str).tolist()
    names = list(df.columns[1:])
    for n in names:
        if pd.api.types.is_numeric_dtype(df[n]):
            continue
        num = pd.to_numeric(df[n], errors="coerce")
        if num.isna().sum() != df[n].isna().sum():
            raise ValueError(f"phenotype column {n!r} holds text values; text coding is not part of this build")
        df[n] = num
    Y = df[names].to_numpy(dtype=np.float64)
    return Phenotypes(iid=iid, names=names, Y=Y)


def read_covar(path) -> Covariates:
    """PLINK-style covariates 'FID IID c1 c2 ...', whitespace separated, no header (io.py:109-120)."""
    import pandas as pd

    df = pd.read_csv(Path(path), sep=r"\s+", header=None)
    cov = df.iloc[:, 2:].to_numpy(dtype=np.float64)
    return Covariates(iid=df.iloc[:, 1].astype(str).tolist(), names=[f"cov_{i}" for i in range(cov.shape[1])], C=cov)


@dataclass
class AlignedDataset:
    """Strain-aligned bundle (io.py:123-140): Z (N, M) raw genotypes, Y (N, P) phenotypes, X (N, C) =
    [intercept | covariates], chrom/pos/snp_id per SNP, optional separate kinship panel Z_kin/chrom_kin."""
    iid: list
    Z: Tensor
    Y: Tensor
    X: Tensor
    chrom: np.ndarray
    pos: np.ndarray
    snp_id: list
    pheno_names: list
    Z_kin: Tensor | None = None
    chrom_kin: np.ndarray | None = None
    n_dropped_kin: int = 0


def align_inputs(geno: Genotypes, pheno: Phenotypes, covar: Covariates | None = None, *,
                 dtype: torch.dtype = torch.float64, kin: Genotypes | None = None) -> AlignedDataset:
    """Strain intersection of geno, pheno, the optional covar and kinship panel, in genotype file order, with
    X = [1 | covariates], as torch tensors of `dtype` (io.py:143-184).  ValueError when no strain is shared."""
    return align_arrays(geno.iid, geno.Z, pheno.iid, pheno.Y,
                        None if covar is None else covar.iid, None if covar is None else covar.C,
                        chrom=geno.chrom, pos=geno.pos, snp_id=geno.sid, pheno_names=pheno.names, dtype=dtype,
                        kin_iid=None if kin is None else kin.iid, Z_kin=None if kin is None else kin.Z,
                        chrom_kin=None if kin is None else kin.chrom)


def align_arrays(geno_iid, Z, pheno_iid, Y, covar_iid=None, C=None, *, chrom, pos=None, snp_id=None,
                 pheno_names=None, dtype=torch.float64, kin_iid=None, Z_kin=None, chrom_kin=None) -> AlignedDataset:
    """align_inputs on plain arrays: strain ID lists with Z (N_g, M), Y (N_p, P), covariates C (N_c, K) and an
    optional kinship panel Z_kin; pos / snp_id / pheno_names default to 0.. and p0.."""
    common = set(geno_iid) & set(pheno_iid)
    if covar_iid is not None:
        common &= set(covar_iid)
    n_dropped_kin = 0
    if kin_iid is not None:
        n_before = len(common)
        common &= set(kin_iid)
        n_dropped_kin = n_before - len(common)
    if not common:
        raise ValueError("no strain overlap between geno / pheno / covar / kinship")
    iid = [s for s in geno_iid if s in common]
    gi = {s: i for i, s in enumerate(geno_iid)}
    pi = {s: i for i, s in enumerate(pheno_iid)}
    Zn = np.asarray(Z)[[gi[s] for s in iid], :]
    Yn = np.asarray(Y, dtype=np.float64)[[pi[s] for s in iid], :]
    if covar_iid is not None:
        ci = {s: i for i, s in enumerate(covar_iid)}
        X = np.concatenate([np.ones((len(iid), 1)), np.asarray(C, dtype=np.float64)[[ci[s] for s in iid], :]], axis=1)
    else:
        X = np.ones((len(iid), 1))
    Zk = None
    if kin_iid is not None:
        ki = {s: i for i, s in enumerate(kin_iid)}
        Zk = torch.from_numpy(np.asarray(Z_kin)[[ki[s] for s in iid], :]).to(dtype)
    M = Zn.shape[1]
    return AlignedDataset(iid=iid, Z=torch.from_numpy(Zn).to(dtype), Y=torch.from_numpy(Yn).to(dtype),
                          X=torch.from_numpy(X).to(dtype), chrom=np.asarray(chrom),
                          pos=np.arange(M) if pos is None else np.asarray(pos),
                          snp_id=[str(i) for i in range(M)] if snp_id is None else list(snp_id),
                          pheno_names=[f"p{i}" for i in range(Yn.shape[1])] if pheno_names is None else list(pheno_names),
                          Z_kin=Zk, chrom_kin=None if chrom_kin is None else np.asarray(chrom_kin),
                          n_dropped_kin=n_dropped_kin)


def group_by_na(Y: Tensor, cols: list[int]) -> list[tuple[list[int], list[int]]]:
    """(keep_rows, cols) per distinct NA pattern, first-seen order (io.py:187-203)."""
    mask = torch.isnan(torch.as_tensor(Y)[:, cols]).cpu().numpy()
    groups: dict[bytes, list[int]] = {}
    for j, _ in enumerate(cols):
        groups.setdefault(np.packbits(mask[:, j]).tobytes(), []).append(j)
    out = []
    for js in groups.values():
        keep = np.flatnonzero(~mask[:, js[0]]).tolist()
        out.append((keep, [cols[j] for j in js]))
    return out


def subset_dataset(data: AlignedDataset, rows: list[int], cols: list[int]) -> AlignedDataset:
    """Strain and phenotype subset; genotypes keep their raw missing calls (io.py:206-218)."""
    r = torch.as_tensor(rows, dtype=torch.long, device=data.Z.device)
    return AlignedDataset(iid=[data.iid[i] for i in rows], Z=data.Z[r], Y=data.Y[r][:, cols], X=data.X[r],
                          chrom=data.chrom, pos=data.pos, snp_id=data.snp_id,
                          pheno_names=[data.pheno_names[c] for c in cols],
                          Z_kin=data.Z_kin[r] if data.Z_kin is not None else None,
                          chrom_kin=data.chrom_kin, n_dropped_kin=data.n_dropped_kin)


def standardise_columns(arr, *, stats: bool = False):
    """Unit standardisation per column on the GPU (io.py:223-254): NaN-aware mean and population std, (x - mean)
    / std, std == 0 -> the column becomes 0, NaN -> 0 after.  Accepts FP64/FP32 (NaN = missing) or int8 (< 0 =
    missing) tensors or numpy arrays; returns the same container type (FP64) on the input's device.  A 1-D input
    is one column.  stats=True also returns (mean, std)."""
    is_np = isinstance(arr, np.ndarray)
    t = torch.from_numpy(np.ascontiguousarray(arr)) if is_np else arr
    home = t.device
    dev = _lib.device()
    one_d = t.ndim == 1
    t2 = t.reshape(-1, 1) if one_d else t
    if t2.dtype == torch.int8:
        src, code = t2.to(dev).contiguous(), 1
    else:
        src, code = t2.to(device=dev, dtype=torch.float64).contiguous(), 0
    n, m = src.shape
    dst = torch.empty(n, m, dtype=torch.float64, device=dev)
    mean = torch.empty(m, dtype=torch.float64, device=dev)
    std = torch.empty(m, dtype=torch.float64, device=dev)
    _lib.call("flmm_standardise", code, src.data_ptr(), n, m, m, dst.data_ptr(), m, mean.data_ptr(), std.data_ptr(),
              _lib.stream())
    out = dst.reshape(-1) if one_d else dst
    out = out.cpu().numpy() if is_np else out.to(home)
    if stats:
        return out, (mean, std)
    return out


def grm(Z: Tensor) -> Tensor:
    """K = Z Z^T / M on pre-standardised Z (io.py:257-265): the FP64 Gram kernel (flmm_gram_f64, FP64 matrix
    cores, deterministic and exactly symmetric) on the GPU; the result comes back on Z's device in Z's dtype."""
    from . import engine

    t = torch.as_tensor(Z)
    home, dt = t.device, t.dtype
    dev = _lib.device()
    K = engine.gram(t.to(device=dev, dtype=torch.float64), prec=engine.prec_code()) / float(t.shape[1])
    return K.to(device=home, dtype=dt)
