"""Write golden input/output vectors for the LMM scan from the reference implementation.

Run in the build container only (the reference tree is not present elsewhere):

    python tests/golden/make_golden.py /root/reference

It decodes the reference's shipped example panel (data/example: 150 strains x 1,500 SNPs on 16 chromosomes,
20 phenotypes, 43 covariates) with a decoder written here, then calls the reference's own
``fasterlmm.core`` / ``fasterlmm.perms`` functions on it and stores inputs and outputs as small .npz files
beside this script.  These pin what only the reference's code can pin: its refined log delta after its own
float ordering, the compat basis with a rank-deficient covariate design, and end-to-end LOCO / permutation
outputs.  Nothing else in the repository reads the reference tree.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def read_bed(prefix: Path) -> tuple[np.ndarray, list[str], np.ndarray]:
    """PLINK 1 SNP-major .bed -> (N, M) float64 dosages, count_A1 convention: 2-bit value 0 -> 2, 1 -> NaN,
    2 -> 1, 3 -> 0 (low bits first within each byte)."""
    fam = [ln.split() for ln in prefix.with_suffix(".fam").read_text().splitlines() if ln.strip()]
    bim = [ln.split() for ln in prefix.with_suffix(".bim").read_text().splitlines() if ln.strip()]
    N, M = len(fam), len(bim)
    raw = np.frombuffer(prefix.with_suffix(".bed").read_bytes(), dtype=np.uint8)
    assert raw[0] == 0x6C and raw[1] == 0x1B and raw[2] == 0x01, "not a SNP-major PLINK 1 bed"
    bps = (N + 3) // 4
    body = raw[3:3 + bps * M].reshape(M, bps)
    codes = np.stack([(body >> (2 * k)) & 3 for k in range(4)], axis=2).reshape(M, bps * 4)[:, :N]
    lut = np.array([2.0, np.nan, 1.0, 0.0])
    Z = lut[codes].T.copy()
    iid = [r[1] for r in fam]
    chrom = np.array([float(r[0]) for r in bim])
    return Z, iid, chrom


def main(ref_root: str) -> None:
    import pandas as pd
    import torch

    sys.path.insert(0, ref_root)
    from fasterlmm import core, io as fio, perms  # the reference implementation, read-only

    torch.set_default_dtype(torch.float64)
    ex = Path(ref_root) / "data" / "example"
    Z_raw, iid, chrom = read_bed(ex / "example")
    ph = pd.read_csv(ex / "example_pheno.tsv", sep="\t")
    assert ph["Strain"].astype(str).tolist() == iid
    Y = ph.iloc[:, 1:].to_numpy(dtype=np.float64)
    cv = pd.read_csv(ex / "example_covar.tab", sep=r"\s+", header=None)
    assert cv.iloc[:, 1].astype(str).tolist() == iid
    covars = cv.iloc[:, 2:].to_numpy(dtype=np.float64)
    N, M = Z_raw.shape
    rng = np.random.default_rng(20240611)
    X_int = np.ones((N, 1))
    X_cov2 = np.concatenate([X_int, rng.normal(size=(N, 2))], axis=1)  # intercept + 2 N(0,1) covariates
    X_full = np.concatenate([X_int, covars], axis=1)  # intercept + 43 one-hot covariates (rank-deficient)
    np.savez_compressed(HERE / "example_inputs.npz", Z_raw=Z_raw.astype(np.int8), Y=Y, chrom=chrom,
                        X_cov2=X_cov2, X_full=X_full, covars=covars)

    T = lambda a: torch.from_numpy(np.ascontiguousarray(a))  # noqa: E731
    Z_std = fio.standardise_columns(T(Z_raw))

    def dump(name, res, extra):
        np.savez_compressed(HERE / name, f=res.f.numpy(), beta=res.beta.numpy(), se=res.se.numpy(),
                            sfve=res.sfve.numpy(), nullh2=res.nullh2.numpy(), max_F=res.max_F.numpy(), **extra)

    def per_chrom_delta(X, Yc, rotate, fit, **kw):
        out = []
        for c in sorted(np.unique(chrom).tolist()):
            K = fio.grm(Z_std[:, T(chrom != c)])
            out.append(fit(rotate(K, T(X), T(Yc)), **kw).numpy())
        return np.stack(out)

    # compat LOCO, three covariate designs; the last one is rank-deficient (one-hot + intercept)
    Y3 = Y[:, :3]
    for tag, X in (("int", X_int), ("cov2", X_cov2), ("full", X_full)):
        res = core.loco_scan_compat(Z_std, T(X), T(Y3), chrom)
        lds = per_chrom_delta(X, Y3, core.fastlmm_compat_rotate, core.fit_delta_grid_compat)
        lds_grid = per_chrom_delta(X, Y3, core.fastlmm_compat_rotate, core.fit_delta_grid_compat, refine=False)
        lds100 = per_chrom_delta(X, Y3, core.fastlmm_compat_rotate, core.fit_delta_grid_compat, n_grid=100)
        dump(f"compat_loco_{tag}.npz", res, {"log_delta": lds, "log_delta_grid": lds_grid,
                                             "log_delta_g100": lds100})
    # compat single-K
    res = core.single_k_scan_compat(Z_std, T(X_cov2), T(Y3))
    dump("compat_single_cov2.npz", res, {})

    # plain GLS LOCO (ML) and the ML / REML per-chromosome log delta
    res = core.loco_scan(Z_std, T(X_cov2), T(Y3), chrom)
    lds_ml = per_chrom_delta(X_cov2, Y3, core.rotate, core.fit_delta_grid)
    lds_reml = per_chrom_delta(X_cov2, Y3, core.rotate, core.fit_delta_grid, reml=True)
    dump("plain_loco_cov2.npz", res, {"log_delta_ml": lds_ml, "log_delta_reml": lds_reml})

    # permutation driver: two phenotypes, four shuffles each; a subset dataset keyed by seed_idx
    data = fio.AlignedDataset(iid=iid, Z=T(Z_raw), Y=T(Y), X=T(X_cov2), chrom=chrom, pos=np.arange(M),
                              snp_id=[str(i) for i in range(M)], pheno_names=list(ph.columns[1:]))
    res, pmf = perms.perm_threshold(data, [0, 1], n_perm=4, seed=19930909, loco=True)
    dump("perm_loco_cov2.npz", res, {"perm_max_F": pmf.numpy()})
    res, pmf = perms.perm_threshold(data, [3], n_perm=3, seed=19930909, loco=False)
    dump("perm_single_cov2.npz", res, {"perm_max_F": pmf.numpy()})
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
