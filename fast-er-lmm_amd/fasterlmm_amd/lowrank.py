"""Exact low-rank spectral LMM on MI355X: the reference's fasterlmm/lowrank.py API (names, dataclasses, signatures)
with the delta search and the Wald scan in the gfx950 kernels.

Reference: fasterlmm/lowrank.py (gbrach/FaST-ER-LMM v1.6.2).

The kinship is a factor G (N, k_in) of pruned markers, K = G G^T / k_in, never formed.  The basis U (N, k) spans
the X-regressed factor; the N - D - k directions outside it see only delta, carried exactly by the off-rank
residuals UUY = Y_r - U U^T Y_r and UUS (lowrank.py:1-10).  ``lowrank_basis(..., rank=r)`` truncates the factor to
its leading r singular directions with the randomised range finder (fasterlmm/extreme.py:23-45) -- the rank-r
kinship of the large-cohort path -- and the exact off-rank algebra then holds for that truncated kinship.

How the kernels carry the off-rank terms (one column p, one SNP s, w = 1/(s_j + delta)):
  * delta search: the compat loss kernels run on the extended spectrum [s, 0] with rotated column
    [U^T y_r ; |UUY|] and log-det spectrum [s, 0 x (Neff - k)], which is the low-rank loss (lowrank.py:128-150)
    term for term; they also produce rWr (on + off rank) and the Q2 planes [w ; 1/delta].
  * scan (flmm_wald_scan_lowrank): sWy = S1.Q1 over K = kpad2 + N with S1 = [U^T s_r ; s_r] and
    Q1 = [w o U^T y_r ; UUY / delta], because UUS^T UUY = s_r^T UUY (UUY is orthogonal to U and X);
    sWs = S2.Q2 over kpad2 only, with S2 = [(U^T s_r)^2 ; |UUS|^2 at entry k].  So the off-rank cross term is
    one K = N contraction on MFMA and the off-rank SNP sum of squares rides as one extra spectral entry.
Where the work runs: the rotations U^T Y_r, U UY and U^T S_r and the range finder's Rayleigh-Ritz product
G^T Q are FP64 on the FP64 matrix cores (flmm_rotate_f64); the range finder's subspace iterations (two K = N or
k_all products per application) run on the MFMA rotation kernel (flmm_rotate) against G's operand planes, packed
once per LOCO run; the small eigendecompositions, CholeskyQR Grams and the C x C covariate algebra are FP64 in
rocSOLVER / rocBLAS through torch.  Scan operand planes are split-FP16 / split-BF16 / FP32 as
``engine.settings.precision`` says, the epilogue FP64.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch
from torch import Tensor

from . import _lib, engine
from .core import ScanResult, _dev, _eigh, _home
from .engine import TILE, PTILE, KCHUNK, rup


@dataclass
class LowRankBasis:
    """Genotype-only part (lowrank.py:25-38): s (k,) ascending kept eigenvalues of the X-regressed kinship, U_eff
    (N, k) orthonormal, X (N, D), Xpinv (D, N), Neff = N - D."""
    s: Tensor
    U_eff: Tensor
    X: Tensor
    Xpinv: Tensor
    Neff: int


@dataclass
class LowRankSpectrum:
    """A LowRankBasis with one column set rotated onto it (lowrank.py:41-56): UY (k, P), UUY (N, P)."""
    s: Tensor
    U_eff: Tensor
    X: Tensor
    Xpinv: Tensor
    Neff: int
    UY: Tensor
    UUY: Tensor


def _regress(X: Tensor, Xpinv: Tensor, A: Tensor) -> Tensor:
    return A - X @ (Xpinv @ A)


def lowrank_basis(G: Tensor, X: Tensor, *, sv_floor: float = 1e-10, rank: int | None = None,
                  n_oversample: int = 20, n_iter: int = 2, seed: int = 19930909) -> LowRankBasis:
    """fastlmm's low-rank decomposition (lowrank.py:59-96): G_ = G - X X^+ G; for k_in < N the eigh of the
    k_in x k_in Gram G_^T G_ (U = G_ V / sv, s = sv^2 / k_in, singular values <= sv_floor dropped), else eigh of
    G_ G_^T / k_in.  ``rank`` keeps only the leading `rank` directions, found by the randomised range finder on
    G_ (extreme.py:23-45; K applied as G_ (G_^T .), never formed), the truncated kinship of the large-cohort
    path.  On the GPU; s comes back ascending."""
    home, (Gd, Xd) = _dev(G, X)
    b = _basis(Gd, Xd, None, sv_floor=sv_floor, rank=rank, n_oversample=n_oversample, n_iter=n_iter, seed=seed)
    return LowRankBasis(s=b.s.to(home), U_eff=b.U_eff.to(home), X=b.X.to(home), Xpinv=b.Xpinv.to(home),
                        Neff=b.Neff)


def _basis(G: Tensor, X: Tensor, keep: Tensor | None, *, sv_floor: float = 1e-10, rank: int | None = None,
           n_oversample: int = 20, n_iter: int = 2, seed: int = 19930909,
           factor: "SplitFactor | None" = None) -> LowRankBasis:
    """lowrank_basis of the columns `keep` (a 0/1 FP64 mask over G's columns, None = all) of the FP64 GPU factor G.
    With a rank the masked factor is never copied (the mask multiplies G^T Y) and the range finder's subspace
    iterations run on the MFMA rotation kernel against G's operand planes (``factor``, packed here if not given);
    the last orthonormalisation and the Rayleigh-Ritz projection are FP64, so U is orthonormal and orthogonal to X
    to FP64 precision."""
    N, k_all = G.shape
    D = X.shape[1]
    Xpinv = torch.linalg.pinv(X)
    k_in = k_all if keep is None else int(keep.sum().item())
    if rank is not None and rank < min(N, k_in):
        s, U = _randomized_factor(G, X, Xpinv, keep, k_in, int(rank), n_oversample, n_iter, seed, factor)
        ok = s.sqrt() * np.sqrt(k_in) > sv_floor
        s, U = s[ok], U[:, ok]
        order = torch.argsort(s)
        return LowRankBasis(s=s[order], U_eff=U[:, order].contiguous(), X=X, Xpinv=Xpinv, Neff=N - D)
    Gk = G if keep is None else G.index_select(1, torch.nonzero(keep).flatten())
    G_ = _regress(X, Xpinv, Gk)
    del Gk
    if k_in < N:
        evals, V = _eigh(G_.T @ G_)
        sv = torch.clamp(evals, min=0.0).sqrt()
        ok = sv > sv_floor
        sv, V = sv[ok], V[:, ok]
        U = G_ @ (V / sv)
        s = sv * sv / k_in
    else:
        evals, Uf = _eigh((G_ @ G_.T) / k_in)
        ok = torch.clamp(evals, min=0.0) > sv_floor * sv_floor / k_in
        s = torch.clamp(evals[ok], min=0.0)
        U = Uf[:, ok].contiguous()
    return LowRankBasis(s=s, U_eff=U, X=X, Xpinv=Xpinv, Neff=N - D)


class SplitFactor:
    """The kinship factor G (N, k_all) as MFMA operand planes in both orientations, packed once and shared by every
    chromosome of a LOCO run: ``gt`` (rows = markers, K = strains) for G^T y and ``g`` (rows = strains, K =
    markers) for G z, in the scan's precision (split-FP16 by default: hi + lo halves with power-of-two row scales,
    three products, FP32 accumulation -- the accuracy of an FP32 GEMM)."""

    def __init__(self, G: Tensor, prec: int):
        N, k_all = G.shape
        G = G.contiguous()
        self.N, self.k_all, self.prec = N, k_all, prec
        self.npad, self.mpad = rup(max(N, 1), TILE), rup(max(k_all, 1), TILE)
        self.gt = engine.pack(prec, G, k_all, N, 1, k_all, self.mpad, self.npad)
        self.g = engine.pack(prec, G, N, k_all, k_all, 1, self.npad, self.mpad)

    def _times(self, a: engine.Planes, rows: int, v: Tensor, kpad: int) -> Tensor:
        """(rows, b) FP64 of a . v for v (K, b) FP64: v^T is packed as the B operand and flmm_rotate forms
        out[i][j] = sum_n a[i][n] v[n][j]."""
        b = v.shape[1]
        v = v.contiguous()
        bp = rup(max(b, 1), TILE)
        vp = engine.pack(self.prec, v, b, v.shape[0], 1, b, bp, kpad)
        out = engine.rotate_uy(a, vp)
        return engine.fp32_plane_to_dense(out, a.rows_pad, bp, rows, b).double()

    def gt_times(self, y: Tensor) -> Tensor:
        """G^T y for y (N, b)."""
        return self._times(self.gt, self.k_all, y, self.npad)

    def g_times(self, z: Tensor) -> Tensor:
        """G z for z (k_all, b)."""
        return self._times(self.g, self.N, z, self.mpad)


def _randomized_factor(G: Tensor, X: Tensor, Xpinv: Tensor, keep: Tensor | None, k_in: int, rank: int,
                       n_oversample: int, n_iter: int, seed: int, factor: SplitFactor | None):
    """Top-`rank` eigenpairs of K_ = G_ G_^T / k_in, G_ = (I - X X^+) G[:, keep], by the randomised range finder
    (Halko, Martinsson & Tropp 2011, alg. 4.4 + 5.3; extreme.py:23-45): Y = K_ Omega and n_iter subspace
    iterations (re-orthonormalised in FP64), each K_ application two MFMA products against G's planes, then an
    orthonormal basis Q of (I - X X^+) Y and the eigh of Q^T K_ Q in FP64."""
    N = G.shape[0]
    target = min(rank + n_oversample, N)
    f = factor if factor is not None else SplitFactor(G, engine.prec_code())
    mask = None if keep is None else keep.unsqueeze(1)

    def apply(Y):
        Z = f.gt_times(_regress(X, Xpinv, Y))
        if mask is not None:
            Z = Z * mask
        return _regress(X, Xpinv, f.g_times(Z)) / k_in

    from .extreme import orthonormal_basis
    gen = torch.Generator(device=G.device).manual_seed(seed)
    Y = apply(torch.randn(N, target, generator=gen, device=G.device, dtype=torch.float32).double())
    for _ in range(n_iter):
        Y = apply(orthonormal_basis(Y))
    Q = orthonormal_basis(_regress(X, Xpinv, Y))
    del Y
    # Rayleigh-Ritz in FP64: Zq = G[:, keep]^T Q on the FP64 matrix cores (Q is orthogonal to X, so G_^T Q = G^T Q)
    cols = None if keep is None else torch.nonzero(keep).flatten()
    n_keep = G.shape[1] if cols is None else cols.numel()
    t = Q.shape[1]
    Zq = engine.rotate_snps_f64(G, cols, n_keep, Q, 0, t, rup(max(n_keep, 1), TILE), rup(t, TILE))[:n_keep, :t]
    B = (Zq.T @ Zq) / k_in
    del Zq
    s_small, U_small = torch.linalg.eigh(0.5 * (B + B.T))
    s_small = torch.clamp(s_small, min=0.0)
    idx = torch.argsort(s_small, descending=True)[:rank]
    return s_small[idx], Q @ U_small[:, idx]


def rotate_phenos(basis: LowRankBasis, Y: Tensor) -> tuple[Tensor, Tensor]:
    """lowrank.py:99-111: Y_r = Y - X X^+ Y, UY = U^T Y_r (k, P), UUY = Y_r - U UY (N, P), FP64 on the GPU: both
    products on the FP64 matrix cores (flmm_rotate_f64)."""
    home, (Yd, Xd, Xp, U) = _dev(Y, basis.X, basis.Xpinv, basis.U_eff)
    Y_r = _regress(Xd, Xp, Yd)
    N, P = Y_r.shape
    k = U.shape[1]
    if k == 0 or P == 0:
        return torch.zeros(k, P, dtype=Y_r.dtype, device=Y_r.device).to(home), Y_r.to(home)
    UY = engine.rotate_snps_f64(Y_r, None, P, U, 0, k, rup(P, TILE), rup(k, TILE))[:P, :k].T.contiguous()
    UUY = Y_r - engine.rotate_snps_f64(U.T.contiguous(), None, N, UY, 0, P, rup(N, TILE), rup(P, TILE))[:N, :P]
    return UY.to(home), UUY.to(home)


def lowrank_rotate(G: Tensor, X: Tensor, Y: Tensor, *, sv_floor: float = 1e-10, rank: int | None = None,
                   **kw) -> LowRankSpectrum:
    """lowrank.py:114-122: basis + one rotation."""
    b = lowrank_basis(G, X, sv_floor=sv_floor, rank=rank, **kw)
    UY, UUY = rotate_phenos(b, Y)
    return LowRankSpectrum(s=b.s, U_eff=b.U_eff, X=b.X, Xpinv=b.Xpinv, Neff=b.Neff, UY=UY, UUY=UUY)


# DELTA SEARCH -------

class _LowRankState:
    """GPU state of one spectrum: the extended spectrum and rotated-column plane that make the compat delta
    kernels evaluate the low-rank loss, and the last fit."""

    def __init__(self, spec: LowRankSpectrum, prec: int):
        _, (s, UY, UUY) = _dev(spec.s, spec.UY, spec.UUY)
        k, P = UY.shape
        self.k, self.P, self.Neff, self.N = k, P, int(spec.Neff), UUY.shape[0]
        self.prec = prec
        dev = s.device
        self.s = torch.cat([s, torch.zeros(1, dtype=torch.float64, device=dev)]).contiguous()  # weight spectrum
        self.s_lam = torch.cat([s, torch.zeros(max(self.Neff - k, 0), dtype=torch.float64, device=dev)]).contiguous()
        self.neff = k + 1
        self.kpad = rup(k + 1, TILE)
        self.p_pad = rup(max(P, 1), PTILE)
        yy_off = (UUY * UUY).sum(dim=0)
        uy_ext = torch.cat([UY, yy_off.sqrt().unsqueeze(0)], dim=0).contiguous()  # (k + 1, P)
        self.uy = engine.pack(_lib.PREC_FP32, uy_ext, P, k + 1, 1, P, self.p_pad, self.kpad).hi
        self.delta = None

    def _loss(self, projected: bool = True):
        return self.s_lam, float(self.Neff)

    def fit(self, n_grid: int, lo: float, hi: float, refine: bool) -> engine.DeltaResult:
        s_lam, nco = self._loss()
        if not refine:
            d = engine.delta_refine(self.prec, _lib.DELTA_GRID, self.uy, self.P, self.p_pad, self.kpad, self.s, s_lam,
                                    nco, n_grid, lo, hi, None)
        elif engine.settings.delta == "learned":
            from . import learned
            fit = learned.refine_learned(self, learned.model(), n_grid, lo, hi, n=self.N, c=self.N - self.Neff)
            self.last_fallback = fit.fallback
            d = fit.delta
        else:
            idx, lam64 = engine.delta_grid(self.uy, self.P, self.p_pad, self.kpad, self.s, s_lam, nco, n_grid, lo,
                                           hi)
            d = engine.delta_refine(self.prec, engine.delta_code(), self.uy, self.P, self.p_pad, self.kpad, self.s,
                                    s_lam, nco, n_grid, lo, hi, idx, lam64=lam64)
        self.delta = d
        return d

    def delta_for(self, log_delta: Tensor) -> engine.DeltaResult:
        d = self.delta
        if d is not None and d.log_delta.data_ptr() == log_delta.data_ptr() and d.q1.prec == self.prec:
            return d
        return engine.delta_refine(self.prec, _lib.DELTA_FIXED, self.uy, self.P, self.p_pad, self.kpad, self.s,
                                   self.s_lam, float(self.Neff), 1, 0.0, 0.0, None, log_delta_in=log_delta)


def _state(spec: LowRankSpectrum) -> _LowRankState:
    st = getattr(spec, "_gpu", None)
    prec = engine.prec_code()
    if st is None or st.prec != prec:
        st = _LowRankState(spec, prec)
        spec._gpu = st
    return st


def _profile_loss_lowrank(log_delta: Tensor, s: Tensor, UY: Tensor, yy_off: Tensor, Neff: int) -> Tensor:
    """(G, P) low-rank -2 log-likelihood on a grid of log delta (lowrank.py:128-150), FP64 on the inputs' device:
    on-rank sum UY^2 / (s + delta) plus the off-rank residual sum of squares yy_off / delta, log det over the k
    kept values plus (Neff - k) log delta.  The scan's delta search evaluates this loss inside the delta kernels
    (on the extended spectrum [s, 0]); this is the reference's helper for its callers."""
    k = s.shape[0]
    delta = log_delta.exp()
    Sd = s.unsqueeze(0) + delta.unsqueeze(1)
    rWr = torch.clamp((1.0 / Sd) @ (UY * UY) + yy_off.unsqueeze(0) / delta.unsqueeze(1),
                      min=1e-300 if s.dtype == torch.float64 else 1e-30)
    logdet = Sd.log().sum(dim=1) + (Neff - k) * delta.log()
    return Neff * (np.log(2.0 * np.pi) + (rWr / Neff).log() + 1.0) + logdet.unsqueeze(1)


def _profile_loss_lowrank_per_pheno(log_delta: Tensor, s: Tensor, UY: Tensor, yy_off: Tensor,
                                    Neff: int) -> Tensor:
    """(P,) low-rank loss with one log delta per column (lowrank.py:153-170)."""
    k = s.shape[0]
    delta = log_delta.exp()
    Sd = s.unsqueeze(0) + delta.unsqueeze(1)
    rWr = torch.clamp(((1.0 / Sd) * UY.T * UY.T).sum(dim=1) + yy_off / delta,
                      min=1e-300 if s.dtype == torch.float64 else 1e-30)
    logdet = Sd.log().sum(dim=1) + (Neff - k) * delta.log()
    return Neff * (np.log(2.0 * np.pi) + (rWr / Neff).log() + 1.0) + logdet


def fit_delta_grid_lowrank(spec: LowRankSpectrum, *, n_grid: int = 64, log_delta_min: float = -10.0,
                           log_delta_max: float = 10.0, refine: bool = True) -> Tensor:
    """lowrank.py:172-208: per-column log delta on the low-rank (ML) loss, grid then refinement in
    [g[i-1], g[i+1]] (Newton to the bracket's minimiser by default, the reference's golden section with
    engine.configure(delta="golden")).  Returns (P,) FP64 on the spectrum's device."""
    st = _state(spec)
    d = st.fit(n_grid, float(log_delta_min), float(log_delta_max), refine)
    return d.log_delta.to(spec.s.device)


# SCAN -------

def _norms(parts: list[Tensor]) -> Tensor:
    """Row norms of the concatenation of several (n_k_i, n_rows) column blocks (each column = one row)."""
    return torch.sqrt(sum((p * p).sum(dim=0) for p in parts))


def pack_concat(prec: int, parts: list[Tensor], rows_pad: int, kpads: list[int]) -> engine.Planes:
    """Operand planes whose K axis is the concatenation of the (n_k_i, n_rows) FP64 blocks `parts` (any strides:
    a transposed view is packed in place), block i occupying kpads[i] (a multiple of 32) entries; split-FP16 rows
    share one scale (from the norm of the whole row), so each block is packed with the combined norm at its K
    offset."""
    n_rows = parts[0].shape[1]
    dev = parts[0].device
    pl = engine.Planes(prec, rows_pad, sum(kpads), dev)
    norm = _norms(parts).contiguous() if pl.prec == _lib.PREC_SPLIT_FP16 else None
    esize = pl.hi.element_size()
    k0 = 0
    st = _lib.stream()
    for part, kp in zip(parts, kpads):
        off = k0 * rows_pad * esize
        hi = pl.hi.data_ptr() + off
        lo = pl.lo.data_ptr() + off if pl.lo is not None else None
        _lib.call("flmm_pack_planes", pl.prec, part.data_ptr(), n_rows, part.shape[0], part.stride(1), part.stride(0),
                  None, _lib.ptr(norm), hi, lo, pl.scale_ptr, rows_pad, kp, st)
        k0 += kp
    return pl


class _ColumnPlanes:
    """Q1 = [w o UY ; UUY / delta] for one delta fit (the column side of the low-rank scan), plus the FP64 pieces
    the re-scoring of large-F tests needs (w, delta, rWr from the FP64 rotated columns)."""

    def __init__(self, st: _LowRankState, spec: LowRankSpectrum, d: engine.DeltaResult):
        _, (s, UY, UUY) = _dev(spec.s, spec.UY, spec.UUY)
        delta = d.log_delta.exp()
        w = 1.0 / (s.unsqueeze(1) + delta.unsqueeze(0))  # (k, P)
        self.npad = rup(st.N, TILE)
        self.kpad2 = st.kpad
        self.q1 = pack_concat(st.prec, [w * UY, UUY / delta.unsqueeze(0)], st.p_pad, [st.kpad, self.npad])
        self.d = d
        self.w, self.delta, self.UY, self.UUY = w, delta, UY, UUY
        self.rWr = torch.clamp((w * UY * UY).sum(dim=0) + (UUY * UUY).sum(dim=0) / delta, min=1e-300)


def snp_wald_scan_lowrank(spec: LowRankSpectrum, log_delta: Tensor, S: Tensor, *, snp_chunk: int | None = None,
                          pheno_chunk: int | None = None, n_real: int | None = None) -> ScanResult:
    """lowrank.py:239-316: Wald F / beta / SE / SFVE / nullh2 of every SNP column of S (N, M; standardised)
    against every column of the spectrum, on- and off-rank terms exact, dof Neff - 1; the first n_real columns
    keep full detail and every column its maximum F.  pheno_chunk is accepted for signature compatibility (all
    columns scan together); SNPs go in blocks of snp_chunk (default 16,384)."""
    st = _state(spec)
    home, (Sd, ldd) = _dev(S, log_delta)
    M = Sd.shape[1]
    n_real = st.P if n_real is None else min(int(n_real), st.P)
    d = st.delta_for(ldd)
    cp = _ColumnPlanes(st, spec, d)
    out = engine.ScanOutputs.allocate(M, n_real, st.P, Sd.device)
    _scan_blocks(st, spec, cp, Sd, torch.arange(M, device=Sd.device), n_real, out, snp_chunk or 16384)
    refresh_real_max(out, n_real)
    return _home(ScanResult(out.f, out.beta, out.se, out.sfve, out.nullh2, out.max_F), home)


def _scan_blocks(st: _LowRankState, spec: LowRankSpectrum, cp: _ColumnPlanes, S: Tensor, row_index: Tensor,
                 n_real: int, out: engine.ScanOutputs, block: int) -> None:
    """Scan the SNP columns S[:, j] (j < S.shape[1]) into output rows row_index[j]."""
    _, (X, Xp, U) = _dev(spec.X, spec.Xpinv, spec.U_eff)
    for b0 in range(0, S.shape[1], block):
        with engine.stage("snps"):
            S_r = _regress(X, Xp, S[:, b0:b0 + block])
            ri = row_index[b0:b0 + block]
            s1, s2, mc, US, ss_off = snp_planes_lowrank(st, U, S_r)
        with engine.stage("scan"):
            wald_scan_lowrank(s1, s2, cp, mc, st, n_real, ri, out)
        del s1, s2
        with engine.stage("rescore"):
            rescore_lowrank(st, cp, S_r, US, ss_off, ri, n_real, out)
        del S_r, US, ss_off


# Real-column tests whose F exceeds this are recomputed in FP64 after the scan: |d log10 p| grows like
# F x (relative F error) / (2 ln 10), so the few very large F (causal hits at large N) would otherwise be the
# first to approach the 1e-3 bar.
RESCORE_F = 30.0


def rescore_lowrank(st: _LowRankState, cp: _ColumnPlanes, S_r: Tensor, US: Tensor, ss_off: Tensor,
                    row_index: Tensor, n_real: int, out: engine.ScanOutputs, chunk: int = 4096) -> None:
    """FP64 re-evaluation of the real-column tests of one SNP block with F > RESCORE_F, from the FP64 operands
    (on-rank US, w, UY; off-rank s_r, UUY, |UUS|^2; rWr): the same Wald arithmetic as the kernel's epilogue
    (lowrank.py:282-301) on exact FP64 sums."""
    if n_real == 0 or S_r.shape[1] == 0:
        return
    F = out.f.index_select(0, row_index)
    mi, pi = torch.nonzero(F > RESCORE_F, as_tuple=True)
    nu = float(st.Neff - 1)
    fl = 1e-300
    for c0 in range(0, mi.numel(), chunk):
        m, p = mi[c0:c0 + chunk], pi[c0:c0 + chunk]
        delta = cp.delta[p]
        w = cp.w[:, p]
        us = US[:, m]
        a = (us * w * cp.UY[:, p]).sum(dim=0) + (S_r[:, m] * cp.UUY[:, p]).sum(dim=0) / delta
        d = torch.clamp((us * us * w).sum(dim=0) + ss_off[m] / delta, min=fl)
        rWr = cp.rWr[p]
        beta = a / d
        var = torch.clamp(rWr - a * a / d, min=fl) / nu / d
        rows = row_index[m]
        out.f[rows, p] = beta * beta / var
        out.beta[rows, p] = beta
        out.se[rows, p] = var.sqrt()
        out.sfve[rows, p] = torch.clamp(a * a / (d * rWr), min=0.0).sqrt()


refresh_real_max = engine.refresh_real_max  # max_F of the real columns from their (re-scored) per-SNP F


def snp_planes_lowrank(st: _LowRankState, U: Tensor, S_r: Tensor):
    """S1 = [U^T s_r ; s_r] (K = kpad2 + N_pad) and S2 = [(U^T s_r)^2 ; |s_r|^2 - |U^T s_r|^2 at entry k]
    (K = kpad2) for the regressed-out SNP block S_r (N, mc)."""
    mc = S_r.shape[1]
    rows_pad = rup(mc, TILE)
    US = engine.rotate_snps_f64(S_r, None, mc, U, 0, st.k, rows_pad, st.kpad)[:mc, :st.k].T  # (k, mc) FP64 view
    us2 = US * US
    ss_off = torch.clamp((S_r * S_r).sum(dim=0) - us2.sum(dim=0), min=0.0)
    s1 = pack_concat(st.prec, [US, S_r], rows_pad, [st.kpad, rup(st.N, TILE)])
    s2 = pack_concat(st.prec, [torch.cat([us2, ss_off.unsqueeze(0)], dim=0)], rows_pad, [st.kpad])
    return s1, s2, mc, US, ss_off


def wald_scan_lowrank(s1: engine.Planes, s2: engine.Planes, cp: _ColumnPlanes, mc: int, st: _LowRankState,
                      n_real: int, row_index: Tensor, out: engine.ScanOutputs) -> None:
    if mc == 0 or st.P == 0:
        return
    d = cp.d
    n_mt = -(-mc // TILE)
    pmax = torch.empty(n_mt * st.p_pad, dtype=torch.float64, device=s1.hi.device)
    o = (out.f, out.beta, out.se, out.sfve, out.nullh2)
    optr = [t.data_ptr() if n_real > 0 else None for t in o]
    (s1h, s1l), (s2h, s2l) = s1.ptrs, s2.ptrs
    (q1h, q1l), (q2h, q2l) = cp.q1.ptrs, d.q2.ptrs
    stream = _lib.stream()
    ev = None
    if engine.scan_timer.enabled:
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        ev[0].record()
    _lib.call("flmm_wald_scan_lowrank", s1.prec, s1h, s1l, s2h, s2l, s1.scale_ptr, s2.scale_ptr, s1.rows_pad,
              q1h, q1l, q2h, q2l, cp.q1.scale_ptr, d.q2.scale_ptr, st.p_pad, s1.kpad, s2.kpad, mc, st.P, n_real,
              d.rWr.data_ptr(), d.h2.data_ptr(), float(st.Neff - 1), row_index.data_ptr(), *optr, pmax.data_ptr(),
              stream)
    if ev is not None:
        ev[1].record()
        # algorithmic flops: on-rank 4 k per test, off-rank cross term 2 N per test
        alg = 4.0 * st.k + 2.0 * st.N
        engine.scan_timer.record(ev, mc * st.P, alg, alg, 2.0 * (s1.kpad + s2.kpad))
    _lib.call("flmm_max_reduce", pmax.data_ptr(), n_mt, st.p_pad, st.P, out.max_F.data_ptr(), stream)


# DRIVERS -------

def loco_scan_lowrank(G: Tensor, g_chrom: np.ndarray, Z: Tensor, z_chrom: np.ndarray, X: Tensor, Y: Tensor, *,
                      n_real: int | None = None, on_chrom=None, rank: int | None = None,
                      snp_chunk: int = 16384) -> ScanResult:
    """lowrank.py:322-360: per chromosome c (sorted labels of z_chrom) the kinship factor drops c's markers, the
    basis, per-column delta and the scan of c's SNPs; max_F is the maximum over chromosomes.  G and Z are
    column-standardised.  ``rank`` truncates each chromosome's kinship (randomised range finder)."""
    from .extreme_scan import _loco_scan_core
    home, (Zd,) = _dev(Z)
    z_chrom = np.asarray(z_chrom)
    chroms = sorted(np.unique(z_chrom).tolist())

    def block_source(c):
        cols = np.flatnonzero(z_chrom == c)
        for s in range(0, len(cols), snp_chunk):
            cidx = cols[s:s + snp_chunk]
            yield Zd.index_select(1, torch.from_numpy(cidx).to(Zd.device)), cidx

    res = _loco_scan_core(chroms, block_source, Zd.shape[1], G, np.asarray(g_chrom), X, Y, n_real=n_real,
                          block_size=snp_chunk, on_chrom=on_chrom, rank=rank)
    return _home(res, home)


def single_k_scan_lowrank(G: Tensor, Z: Tensor, X: Tensor, Y: Tensor, *, n_real: int | None = None,
                          rank: int | None = None) -> ScanResult:
    """lowrank.py:363-373: one kinship factor, one rotation + delta fit, one scan over all of Z."""
    spec = lowrank_rotate(G, X, Y, rank=rank)
    return snp_wald_scan_lowrank(spec, fit_delta_grid_lowrank(spec), Z, n_real=n_real)
