"""Seeded synthetic GWAS panels for the scan's workloads.

Genotypes follow the Balding-Nichols model: ancestral frequency p ~ U(maf_lo, maf_hi) per SNP, subpopulation
frequency ~ Beta(p (1 - F) / F, (1 - p)(1 - F) / F) with F = Fst, diploid dosages ~ Binomial(2, p_k), stored as
int8 0/1/2.  SNPs are split evenly over chromosomes 1..n_chrom (sorted).  Each phenotype has heritability
h2 ~ U(h2_lo, h2_hi) from n_causal SNPs with N(0, 1) effects on standardised genotypes, plus Gaussian noise;
optionally masked by seeded missing-at-random NA groups (na_mask) and rank-inverse-normal transformed with Blom's c = 3/8 and average ties (the reference's RINT,
fasterlmm/normalize.py:21-52).  Covariates are an intercept plus N(0, 1) columns.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.stats import norm, rankdata


@dataclass
class Panel:
    G: np.ndarray        # (N, M) int8 dosages
    chrom: np.ndarray    # (M,) int chromosome labels, sorted
    X: np.ndarray        # (N, C) covariates, intercept first
    Y: np.ndarray        # (N, P) phenotypes
    h2: np.ndarray       # (P,) simulated heritabilities
    causal: np.ndarray   # (P, n_causal) causal SNP indices


CONFIGS = {
    # config 1: the reference's CPU-runnable example shape
    1: dict(N=150, M=1500, n_chrom=16, P=20, C=3, maf=(0.05, 0.5), h2=(0.1, 0.8), rint=False, n_perm=100),
    # config 2: the headline single-GPU scan
    2: dict(N=1000, M=100_000, n_chrom=22, P=6484, C=4, maf=(0.01, 0.5), h2=(0.0, 0.9), rint=True, n_perm=100),
    # config 3: uneven batching -- 64 missingness groups, retained N ~ U(3000, 5000)
    3: dict(N=5000, M=200_000, n_chrom=22, P=2000, C=6, maf=(0.01, 0.5), h2=(0.0, 0.9), rint=True, n_perm=100,
            na_groups=64, n_keep=(3000, 5000), sampler="bernoulli2"),
}


def rint_columns(Y: np.ndarray, c: float = 3.0 / 8.0) -> np.ndarray:
    """Per-column Blom RINT: qnorm((rank - c) / (n - 2c + 1)), average ties, NaN kept (normalize.py:21-52)."""
    Y = np.asarray(Y, dtype=np.float64)
    out = np.full_like(Y, np.nan)
    for j in range(Y.shape[1]):
        m = ~np.isnan(Y[:, j])
        n = int(m.sum())
        if n:
            out[m, j] = norm.ppf((rankdata(Y[m, j], method="average") - c) / (n - 2 * c + 1))
    return out


def balding_nichols(rng: np.random.Generator, N: int, M: int, *, n_pop: int = 4, fst: float = 0.1,
                    maf=(0.05, 0.5), block: int = 8192, sampler: str = "binomial") -> np.ndarray:
    """sampler "binomial" draws each dosage with rng.binomial(2, p); "bernoulli2" draws the same Binomial(2, p)
    law as the sum of two float32-uniform Bernoulli(p) trials, several times faster for large panels."""
    G = np.empty((N, M), dtype=np.int8)
    pop = np.repeat(np.arange(n_pop), -(-N // n_pop))[:N]
    for j0 in range(0, M, block):
        j1 = min(M, j0 + block)
        p = rng.uniform(maf[0], maf[1], size=j1 - j0)
        a = p * (1 - fst) / fst
        b = (1 - p) * (1 - fst) / fst
        pk = rng.beta(a[None, :], b[None, :], size=(n_pop, j1 - j0))
        if sampler == "bernoulli2":
            pf = pk[pop].astype(np.float32)
            G[:, j0:j1] = (rng.random((N, j1 - j0), dtype=np.float32) < pf).astype(np.int8)
            G[:, j0:j1] += (rng.random((N, j1 - j0), dtype=np.float32) < pf).astype(np.int8)
        else:
            G[:, j0:j1] = rng.binomial(2, pk[pop]).astype(np.int8)
    # a column with no variation scans as zero; keep every SNP polymorphic
    mono = np.flatnonzero(G.min(axis=0) == G.max(axis=0))
    for j in mono:
        G[rng.integers(0, N), j] = 1 if G[0, j] != 1 else 0
    return G


def na_mask(rng: np.random.Generator, N: int, P: int, n_groups: int, n_keep) -> np.ndarray:
    """(N, P) missing-at-random mask of n_groups NA patterns: group g holds the columns j with j % n_groups == g
    and keeps a uniformly drawn set of n_g ~ U{n_keep[0], ..., n_keep[1]} strains (n_g <= N)."""
    mask = np.ones((N, P), dtype=bool)
    for g in range(n_groups):
        n_g = min(N, int(rng.integers(n_keep[0], n_keep[1] + 1)))
        keep = rng.choice(N, size=n_g, replace=False)
        cols = np.arange(g, P, n_groups)
        mask[np.ix_(keep, cols)] = False
    return mask


def make_panel(config: int | dict, seed: int = 19930909, *, n_causal: int = 10) -> Panel:
    cfg = CONFIGS[config] if isinstance(config, int) else config
    rng = np.random.default_rng(seed)
    N, M, P, C = cfg["N"], cfg["M"], cfg["P"], cfg["C"]
    G = balding_nichols(rng, N, M, maf=cfg["maf"], sampler=cfg.get("sampler", "binomial"))
    per = -(-M // cfg["n_chrom"])
    chrom = (np.arange(M) // per + 1).astype(np.int64)
    X = np.concatenate([np.ones((N, 1)), rng.standard_normal((N, C - 1))], axis=1)
    h2 = rng.uniform(cfg["h2"][0], cfg["h2"][1], size=P)
    causal = np.stack([rng.choice(M, size=n_causal, replace=False) for _ in range(P)])
    Y = np.empty((N, P))
    for p0 in range(0, P, 1024):
        p1 = min(P, p0 + 1024)
        idx = causal[p0:p1]  # (b, n_causal)
        Gc = G[:, idx.reshape(-1)].astype(np.float64).reshape(N, p1 - p0, n_causal)
        mu = Gc.mean(axis=0, keepdims=True)
        sd = Gc.std(axis=0, keepdims=True)
        sd[sd == 0] = 1.0
        g = (((Gc - mu) / sd) * rng.standard_normal((1, p1 - p0, n_causal))).sum(axis=2)
        g = (g - g.mean(axis=0)) / np.maximum(g.std(axis=0), 1e-12)
        hh = h2[p0:p1]
        Y[:, p0:p1] = np.sqrt(hh) * g + np.sqrt(1 - hh) * rng.standard_normal((N, p1 - p0))
    if cfg.get("na_groups"):
        Y[na_mask(rng, N, P, cfg["na_groups"], cfg["n_keep"])] = np.nan
    if cfg.get("rint"):
        Y = rint_columns(Y)
    return Panel(G=G, chrom=chrom, X=X, Y=Y, h2=h2, causal=causal)


# config 5: the large-cohort low-rank workload, generated on the GPU
CONFIGS[5] = dict(N=50_000, M=500_000, n_chrom=22, P=1000, C=10, maf=(0.01, 0.5), h2=(0.0, 0.9), rint=False,
                  n_perm=20, n_kin=50_000, rank=2000, block=16384)


@dataclass
class GpuPanel:
    G: "torch.Tensor"      # (N, M) int8 dosages on the GPU
    chrom: np.ndarray      # (M,) chromosome labels, sorted
    kin_idx: np.ndarray    # kinship marker columns (an even stride across the genome)
    X: "torch.Tensor"      # (N, C) FP64 covariates, intercept first
    Y: "torch.Tensor"      # (N, P) FP64 phenotypes


def make_panel_gpu(cfg: dict, seed: int = 19930909, *, n_causal: int = 10, block: int = 16384,
                   device=None) -> GpuPanel:
    """The same model as make_panel (Balding-Nichols genotypes over 4 subpopulations, Fst = 0.1, phenotypes with
    n_causal standardised causal SNPs at heritability h2, intercept + N(0, 1) covariates) drawn with torch's GPU
    generator, for panels too large to draw on the host: dosages are sums of two Bernoulli(p_k) draws."""
    import torch
    dev = torch.device("cuda") if device is None else torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    N, M, P, C = cfg["N"], cfg["M"], cfg["P"], cfg["C"]
    fst, n_pop = 0.1, 4
    pop = torch.arange(N, device=dev) * n_pop // N
    G = torch.empty((N, M), dtype=torch.int8, device=dev)
    for j0 in range(0, M, block):
        j1 = min(M, j0 + block)
        p = torch.rand(j1 - j0, generator=gen, device=dev, dtype=torch.float64) * (cfg["maf"][1] - cfg["maf"][0]) \
            + cfg["maf"][0]
        a = p * (1 - fst) / fst
        b = (1 - p) * (1 - fst) / fst
        ga = torch._standard_gamma(a.unsqueeze(0).expand(n_pop, -1).contiguous(), generator=gen)
        gb = torch._standard_gamma(b.unsqueeze(0).expand(n_pop, -1).contiguous(), generator=gen)
        pk = (ga / (ga + gb)).float()[pop]  # (N, b) subpopulation frequency of each strain
        g = (torch.rand(pk.shape, generator=gen, device=dev) < pk).to(torch.int8)
        g += (torch.rand(pk.shape, generator=gen, device=dev) < pk).to(torch.int8)
        G[:, j0:j1] = g
        del pk, g
    per = -(-M // cfg["n_chrom"])
    chrom = (np.arange(M) // per + 1).astype(np.int64)
    n_kin = min(cfg.get("n_kin", M), M)
    kin_idx = np.unique(np.linspace(0, M - 1, n_kin).round().astype(np.int64))
    X = torch.cat([torch.ones(N, 1, device=dev, dtype=torch.float64),
                   torch.randn(N, C - 1, generator=gen, device=dev, dtype=torch.float64)], dim=1)
    h2 = torch.rand(P, generator=gen, device=dev, dtype=torch.float64) * (cfg["h2"][1] - cfg["h2"][0]) + cfg["h2"][0]
    causal = torch.randint(0, M, (P, n_causal), generator=gen, device=dev)
    Y = torch.empty((N, P), dtype=torch.float64, device=dev)
    for p0 in range(0, P, 256):
        p1 = min(P, p0 + 256)
        Gc = G[:, causal[p0:p1].reshape(-1)].double().reshape(N, p1 - p0, n_causal)
        Gc = (Gc - Gc.mean(0, keepdim=True)) / Gc.std(0, unbiased=False, keepdim=True).clamp(min=1e-12)
        g = (Gc * torch.randn(1, p1 - p0, n_causal, generator=gen, device=dev, dtype=torch.float64)).sum(2)
        g = (g - g.mean(0)) / g.std(0, unbiased=False).clamp(min=1e-12)
        e = torch.randn(N, p1 - p0, generator=gen, device=dev, dtype=torch.float64)
        Y[:, p0:p1] = h2[p0:p1].sqrt() * g + (1 - h2[p0:p1]).sqrt() * e
    return GpuPanel(G=G, chrom=chrom, kin_idx=kin_idx, X=X, Y=Y)
