"""Drop-in spectral LMM core on MI355X: the reference's fasterlmm/core.py API (names, signatures, dataclasses,
error behaviour) with the rotations, the delta search and the Wald scan running in the gfx950 kernels.

Reference: fasterlmm/core.py (gbrach/FaST-ER-LMM v1.6.2).

* Inputs may live on the CPU or the GPU; the work always runs on the GPU and the outputs come back on the
  device of the input, in FP64, as in the reference (core.py:322-326, 701-706).  There is no CPU fallback:
  without a ROCm GPU or without the built library every entry point raises.
* ``snp_chunk`` / ``pheno_chunk`` are accepted for signature compatibility; tiling is internal (128 x 128 MFMA
  tiles) and, as in the reference (its chunking-invariance tests), invisible in the results.
* Precision: the column rotation and the two scan products run as split-FP16 MFMA by default (hi + lo IEEE
  half on power-of-two-scaled rows, ~22 significant bits, three products, FP32 accumulation); split-BF16
  (``engine.configure(precision="split_bf16")``, ~16 bits) and FP32 MFMA (``precision="fp32"``) are selectable.
  The kinship Gram, the SNP rotation, the Wald epilogue, the delta search, the projections and the
  eigendecompositions are FP64.
* The plain GLS path needs a full-column-rank covariate design X and raises ValueError otherwise: the
  reference's Cholesky of X~^T V^-1 X~ fails on a singular Gram and it substitutes an identity factor and an
  infinite loss (core.py:166-171, 336-340), which leaves its outputs decided by rounding.  The compat path
  handles a rank-deficient X through X^+ as the reference does.
* Delta refinement defaults to safeguarded Newton to the local minimiser inside the reference's golden-section
  bracket; ``engine.configure(delta="golden")`` replays the reference's 50 golden-section steps instead.

One projected basis serves both formulations.  For an orthonormal basis B of null(X^T) that diagonalises the
projected kinship, B^T (K + delta I) B = diag(s' + delta), the GLS Wald statistic of a SNP with the covariates X
as nuisance equals the covariate-free weighted regression of B^T y on B^T s with weights 1 / (s' + delta):
GLS's M = V^-1 - V^-1 X (X^T V^-1 X)^-1 X^T V^-1 equals B (B^T V B)^-1 B^T.
  - compat path (core.py:454-740): B = the kept eigenvectors of (I - X X^+)(K + I)(I - X X^+), the reference's
    own construction.
  - plain GLS path (core.py:89-451): the same construction on K + 1e-6 I, so every test needs 2 MFMA products
    instead of C + 2 and the C x C Schur complement (core.py:351-355) is carried out once per chromosome in
    FP64 by the basis change instead of per test.  The ML loss keeps its full log det V term (the spectrum of
    K + 1e-6 I) and N; the REML loss equals the projected loss up to a constant.
The test SNPs' OLS projection S - X X^+ S (core.py:628) is not formed: B^T X = 0, so B^T (S - X X^+ S) = B^T S.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch
from torch import Tensor

from . import _lib, engine
from .engine import TILE, PTILE, KCHUNK, rup


# SCAN ARRAYS -------

@dataclass
class Spectrum:
    """K's eigendecomposition + rotated design matrices (core.py:19-31)."""
    s: Tensor
    U: Tensor
    X_rot: Tensor
    Y_rot: Tensor


@dataclass
class ScanResult:
    """Per-(SNP, real column) Wald outputs and the per-column genome max F (core.py:34-48)."""
    f: Tensor
    beta: Tensor
    se: Tensor
    sfve: Tensor
    nullh2: Tensor
    max_F: Tensor


@dataclass
class CompatSpectrum:
    """Projected eigendecomposition of the fastlmm-compat path (core.py:466-484)."""
    s: Tensor
    U_eff: Tensor
    X: Tensor
    Xpinv: Tensor
    UY: Tensor


_FIELDS = ("f", "beta", "se", "sfve", "nullh2", "max_F")


def _pos_floor(dtype: torch.dtype) -> float:
    """Clamp floor per dtype (core.py:80-86)."""
    return 1e-300 if dtype == torch.float64 else 1e-30


def _dev(*ts):
    """Move tensors to the GPU as FP64; return (device of the first input, moved tensors)."""
    dev = _lib.device()
    home = ts[0].device if isinstance(ts[0], Tensor) else torch.device("cpu")
    out = [None if t is None else torch.as_tensor(t).to(device=dev, dtype=torch.float64) for t in ts]
    return home, out


def _home(res: ScanResult, home: torch.device) -> ScanResult:
    if home.type == "cuda":
        return res
    return ScanResult(*(getattr(res, k).to(home) for k in _FIELDS))


def _eigh(K: Tensor) -> tuple[Tensor, Tensor]:
    """FP64 symmetric eigendecomposition on K's device (rocSOLVER through torch on the GPU), ascending
    (core.py:54-77).  An out-of-memory error retries on the host LAPACK as the reference's fallback does (its
    workspace is several times K); other errors propagate."""
    try:
        return torch.linalg.eigh(K)
    except RuntimeError as e:
        if "out of memory" not in str(e).lower():
            raise
        dev = K.device
        Kc = K.detach().cpu()
        del K
        torch.cuda.empty_cache()
        s, U = torch.linalg.eigh(Kc)
        return s.to(dev), U.to(dev)


_eigh_with_cpu_fallback = _eigh  # the reference's name (core.py:54), imported by its tests


def eigendecompose(K: Tensor, jitter: float = 1e-6) -> tuple[Tensor, Tensor]:
    """eigh(K + jitter I) with eigenvalues clamped at 0 (core.py:89-102)."""
    home, (Kd,) = _dev(K)
    N = Kd.shape[0]
    s, U = _eigh(Kd + jitter * torch.eye(N, dtype=Kd.dtype, device=Kd.device))
    s = torch.clamp(s, min=0.0)
    return s.to(home), U.to(home)


# PROJECTED BASIS ON THE GPU -------

# Real-column tests with F above this are re-evaluated in FP64 (_Basis.rescore).  The scan's FP32-accumulated sums
# carry ~1e-7 relative error, and rss = rWr - sWy^2 / sWs amplifies it by ~1 + F / nu, while |d log10 p| grows like
# F x (relative F error) / (2 ln 10): below F = 30 the error stays near 1e-6, far inside the 1e-3 bar.
RESCORE_F = 30.0
EXACT_COLS = 4096  # leading columns whose rotation rotate_columns takes from the FP64 product by default


class _Basis:
    """A projected basis B (columns col0 .. col0 + n_basis - 1 of the (N, L) FP64 matrix `mat`) with its weight
    spectrum s (1 / (s + delta) per basis vector), held as MFMA planes of B^T, plus the rotated scan columns.

    nu is the Wald dof.  Losses: the projected (compat / REML) loss has coefficient n_basis and log-det spectrum
    s; the GLS ML loss has coefficient N and log-det spectrum s_ml (the full spectrum of K + jitter I)."""

    def __init__(self, s: Tensor, mat: Tensor, col0: int, n_basis: int, prec: int, nu: float,
                 s_ml: Tensor | None = None):
        N, L = mat.shape
        self.N, self.neff = N, n_basis
        self.npad = rup(N, KCHUNK)
        self.kpad = rup(max(n_basis, 1), TILE)
        self.prec = prec
        self.nu = float(nu)
        self.s = s.contiguous()
        self.s_range = tuple(torch.stack(torch.aminmax(self.s)).tolist()) if n_basis > 0 else (0.0, 0.0)
        self.s_ml = None if s_ml is None else s_ml.contiguous()
        mat = mat.contiguous()
        # B^T planes: row j = basis vector j, element k = mat[k, col0 + j]
        self.bt = engine.pack(prec, mat, n_basis, N, 1, L, self.kpad, self.npad, offset=col0)
        self.mat, self.col0 = mat, col0  # the FP64 basis: SNP rotation, exact column rows, re-scoring
        self.y_cols = None  # FP64 scan columns (at least the real ones), for the FP64 re-scoring
        self._u_pad = None
        self._uy64 = {}
        self.uy = None  # fp32 plane of B^T Y
        self.col_scale = None  # split-FP16 columns inside span(B): the Y planes' row scales also bound B^T y
        self.P = 0
        self.p_pad = 0
        self.delta = None  # the DeltaResult of the last fit (its log_delta tensor identifies it)

    def rotate_columns(self, Y: Tensor, n_exact: int | None = None, *, in_span: bool = False) -> None:
        """UY plane of B^T Y for Y (N, P) FP64; its first n_exact columns (default: all, up to EXACT_COLS) from the
        FP64 product.  in_span: the columns lie in span(B) (regressed on X, B a complete basis of null(X^T)), so
        |B^T y| = |y| and the planes' row scales serve the split delta grid."""
        N, P = Y.shape
        self.P, self.p_pad = P, rup(max(P, 1), PTILE)
        ypl = engine.pack(self.prec, Y.contiguous(), P, N, 1, P, self.p_pad, self.npad)
        self.uy = engine.rotate_uy(ypl, self.bt)
        self.col_scale = ypl.scale if in_span else None
        self.y_cols, self._uy64 = Y, {}
        self._exact_rows(Y[:, : min(P, EXACT_COLS if n_exact is None else int(n_exact))])

    def use_column_planes(self, ypl: engine.Planes, P: int, y_real: Tensor | None = None, *,
                          in_span: bool = False) -> None:
        """UY plane from pre-packed column planes (shared by every chromosome of a LOCO scan); y_real: the FP64
        real columns (N, n_real), whose rows come from the FP64 product and which the re-scoring reads; in_span as
        rotate_columns."""
        self.P, self.p_pad = P, ypl.rows_pad
        self.uy = engine.rotate_uy(ypl, self.bt)
        self.col_scale = ypl.scale if in_span else None
        self.y_cols, self._uy64 = y_real, {}
        if y_real is not None:
            self._exact_rows(y_real)

    def _exact_rows(self, y: Tensor) -> None:
        """Rows 0 .. n-1 of the UY plane (the leading, real columns y (N, n) FP64) from the FP64 product B^T y,
        rounded once to FP32.  The split rotation's error is ~2^-22 of |y| in every basis direction; the scan weighs
        direction k by 1 / (s_k + delta), so where a column's delta and some eigenvalues are small (h2 near 1,
        delta ~ 1e-4 at the grid's lower edge) that error would reach beta and the delta fit amplified ~1e4
        times, while the FP64 product leaves each entry's own FP32 rounding only."""
        n = y.shape[1]
        if n == 0 or self.neff == 0:
            return
        # B^T y on the FP64 matrix cores (flmm_rotate_f64): rows i < n, entries j < neff, zero beyond
        UY = engine.rotate_snps_f64(y, None, n, self.mat, self.col0, self.neff, rup(n, TILE), self.kpad)[:n]
        self._uy64[n] = UY  # the re-scoring's FP64 operand
        nch = self.kpad // KCHUNK
        self.uy.view(nch, self.p_pad, KCHUNK)[:, :n, :] = UY.float().view(n, nch, KCHUNK).permute(1, 0, 2)

    def _rotate_f64(self, cols: Tensor) -> Tensor:
        """FP64 B^T c for the columns c of cols (N, n): rows (rup(n, 64), kpad), row i = column i rotated."""
        if self._u_pad is None:
            self._u_pad = torch.zeros(self.npad, self.kpad, dtype=torch.float64, device=cols.device)
            self._u_pad[: self.N, : self.neff] = self.mat[:, self.col0:self.col0 + self.neff]
        n = cols.shape[1]
        rows = rup(max(n, 1), 64)
        a = torch.zeros(rows, self.npad, dtype=torch.float64, device=cols.device)
        a[:n, : self.N] = cols.T
        c = torch.empty(rows, self.kpad, dtype=torch.float64, device=cols.device)
        _lib.call("flmm_gemm_f64", a.data_ptr(), rows, self.npad, self._u_pad.data_ptr(), self.kpad, c.data_ptr(),
                  _lib.stream())
        return c

    def rescore(self, columns, local: Tensor, row_index: Tensor, n_real: int, out: engine.ScanOutputs,
                d: engine.DeltaResult, us: Tensor | None = None, row_range: tuple[int, int] | None = None) -> int:
        """FP64 re-evaluation (flmm_rescore_f64) of one scanned SNP block's real-column tests with F > RESCORE_F.
        Block position i holds SNP column local[i] of columns(idx) -> (N, len(idx)) FP64 and output row
        row_index[i]; us, if given, the block's FP64 rotated SNPs (row i = block position i), which then need no
        second rotation; row_range (start, stop), if given, says row_index is start, start + 1, ..., stop - 1.
        Returns the number of re-scored tests, or -1 where both us and row_range are given (the count then stays on
        the device: the host does not wait for the scan)."""
        if n_real == 0 or self.y_cols is None or row_index.numel() == 0:
            return 0
        UY = self._uy64_for(n_real)
        if us is not None and row_range is not None:
            # consecutive output rows and the block's FP64 rotated SNPs at hand: select into lists that hold every
            # test of the block and re-score on the device's count, with no host read (-1: count not read)
            F = out.f[row_range[0]:row_range[1]]
            cap = F.shape[0] * n_real
            mi = torch.empty(cap, dtype=torch.int64, device=F.device)
            pi = torch.empty(cap, dtype=torch.int64, device=F.device)
            cnt = torch.zeros(1, dtype=torch.int64, device=F.device)
            st = _lib.stream()
            _lib.call("flmm_select_above", F.data_ptr(), F.shape[0], n_real, F.stride(0), float(RESCORE_F),
                      mi.data_ptr(), pi.data_ptr(), cap, cnt.data_ptr(), st)
            _lib.call("flmm_rescore_f64", us.data_ptr(), UY.data_ptr(), self.kpad, self.s.data_ptr(), self.neff,
                      d.log_delta.data_ptr(), mi.data_ptr(), pi.data_ptr(), None, int(row_range[0]), cap,
                      cnt.data_ptr(), float(self.nu), out.f.data_ptr(), out.beta.data_ptr(), out.se.data_ptr(),
                      out.sfve.data_ptr(), out.f.shape[1], st)
            return -1
        if row_range is not None:  # consecutive output rows: a view instead of a gathered copy
            F = out.f[row_range[0]:row_range[1]]
        else:
            F = out.f.index_select(0, row_index)
        mi, pi = _select_above(F, RESCORE_F)  # block positions and columns of the tests with F above the bar
        n = int(mi.numel())
        if n == 0:
            return 0
        if us is not None:
            US, inv = us, mi
        else:
            um, inv = torch.unique(mi, return_inverse=True)
            inv = inv.contiguous()
            US = self._rotate_f64(columns(local.index_select(0, um)))
        _lib.call("flmm_rescore_f64", US.data_ptr(), UY.data_ptr(), self.kpad, self.s.data_ptr(), self.neff,
                  d.log_delta.data_ptr(), inv.data_ptr(), pi.data_ptr(), row_index.index_select(0, mi).data_ptr(), 0,
                  n, None, float(self.nu), out.f.data_ptr(), out.beta.data_ptr(), out.se.data_ptr(),
                  out.sfve.data_ptr(), out.f.shape[1], _lib.stream())
        return n

    def _uy64_for(self, n_real: int) -> Tensor:
        """The first n_real columns' FP64 rotation (cached: it does not depend on the SNP block, so one FP64
        product keeps results block-invariant)."""
        UY = self._uy64.get(n_real)
        wider = [k for k in self._uy64 if k > n_real]
        if UY is None and wider:
            UY = self._uy64[min(wider)][:n_real]
        if UY is None:
            yc = self.y_cols[:, :n_real].contiguous()
            UY = engine.rotate_snps_f64(yc, None, n_real, self.mat, self.col0, self.neff, rup(n_real, TILE),
                                        self.kpad)[:n_real]
            self._uy64[n_real] = UY
        return UY

    def uy_dense(self) -> Tensor:
        return engine.fp32_plane_to_dense(self.uy, self.p_pad, self.kpad, self.P, self.neff).T.double()

    def _loss(self, reml_or_compat: bool):
        if reml_or_compat or self.s_ml is None:
            return self.s, float(self.neff)
        return self.s_ml, float(self.N)

    def factored(self, lo: float, hi: float) -> engine.Factored | None:
        """Factored sWs weights for log delta in [lo, hi] (engine.settings.weights: "auto" = for the split
        precisions), or None for the exact weights."""
        w = engine.settings.weights
        if w not in ("auto", "factored", "exact"):
            raise ValueError(f"unknown weights setting {w!r}")
        if w == "exact" or (w == "auto" and self.prec == _lib.PREC_FP32) or self.neff == 0:
            return None
        if not (math.isfinite(lo) and math.isfinite(hi)):
            return None
        return engine.factored_weights(self.s, self.s_range, self.kpad, lo, hi)

    def fit(self, n_grid: int, lo: float, hi: float, refine: bool, *, projected_loss: bool = True,
            n_real: int | None = None):
        """Delta search of every column; n_real (the leading real columns), when given, lets the columns of the
        certified screen get their hi planes only."""
        s_lam, nco = self._loss(projected_loss)
        fac = self.factored(lo, hi)  # every refined or grid log delta lies in [lo, hi]
        hf = engine.screen_start(self.prec, n_real, self.P) if self.neff > 0 else None
        if refine and engine.settings.delta == "learned":
            from . import learned
            fit = learned.refine_learned(self, learned.model(), n_grid, lo, hi, n=self.N, c=self.N - self.neff,
                                         projected_loss=projected_loss, hi_from=hf)
            if fac is not None:
                engine.attach_factored(self.prec, fit.delta, fac)
            self.last_fallback = fit.fallback
            return fit.log_delta, fit.delta
        if not refine:
            d = engine.delta_refine(self.prec, _lib.DELTA_GRID, self.uy, self.P, self.p_pad, self.kpad, self.s, s_lam,
                                    nco, n_grid, lo, hi, None, fac=fac, hi_from=hf)
            self.delta = d
            return d.log_delta, d
        idx, lam64 = engine.delta_grid(self.uy, self.P, self.p_pad, self.kpad, self.s, s_lam, nco, n_grid, lo, hi,
                                       col_scale=self.col_scale, s_min=self.s_range[0])
        d = engine.delta_refine(self.prec, engine.delta_code(), self.uy, self.P, self.p_pad, self.kpad, self.s, s_lam,
                                nco, n_grid, lo, hi, idx, fac=fac, lam64=lam64, hi_from=hf)
        self.delta = d
        return d.log_delta, d

    def delta_for(self, log_delta: Tensor, n_real: int | None = None) -> engine.DeltaResult:
        """The column planes for given log delta values (n_real: as in fit); the screened columns' bucketed layout
        covers the values' range."""
        d = self.delta
        if d is not None and d.log_delta.data_ptr() == log_delta.data_ptr() and d.q1.prec == self.prec:
            return d
        fac = None
        t_range = None
        if log_delta.numel() > 0:
            t_range = tuple(float(v) for v in torch.aminmax(log_delta.double()))
            fac = self.factored(*t_range)
        hf = engine.screen_start(self.prec, n_real, self.P) if self.neff > 0 else None
        return engine.delta_refine(self.prec, _lib.DELTA_FIXED, self.uy, self.P, self.p_pad, self.kpad, self.s,
                                   self.s, float(self.neff), 1, 0.0, 0.0, None, log_delta_in=log_delta, fac=fac,
                                   hi_from=hf, t_range=t_range)

    def scan(self, Z: Tensor, row_index: Tensor, n_real: int, out: engine.ScanOutputs, d: engine.DeltaResult,
             block: int = 8192, after_launch=None, first_row: int | None = None) -> None:
        """Scan the SNP columns Z[:, row_index] of Z (N, M_total) FP64 into out.  after_launch() is called once,
        after the first block's scan is queued and before the host waits for it (the FP64 re-scoring's selection
        reads F back).  first_row, if given, says row_index is first_row, first_row + 1, ..."""
        Mc = row_index.numel()
        screen = engine.screen_applies(self.prec) and self.P > n_real and self.neff > 0
        f64 = engine.settings.snp_rotation == "fp64"
        for b0 in range(0, Mc, block):
            ri = row_index[b0:b0 + block]
            scr = us = None
            if f64:  # the block's FP64 rotated SNPs also serve the screen and the re-scoring
                s1, s2, us = self.snp_planes(Z, ri, d, keep_us=True)
            else:
                s1, s2 = self.snp_planes(Z, ri, d)
            if screen:
                scr = engine.ScreenOperands(us, self.uy, self.s, self.s_range[0], self.s_range[1], self.neff)
            engine.wald_scan(s1, s2, d, ri.numel(), self.P, n_real, self.nu, ri, out, screen=scr)
            del s1, s2, scr
            if after_launch is not None:
                after_launch()
                after_launch = None
            rr = None if first_row is None else (first_row + b0, first_row + b0 + ri.numel())
            self.rescore(lambda idx: Z.index_select(1, idx), ri, ri, n_real, out, d, us=us, row_range=rr)
            del us
        if after_launch is not None:
            after_launch()

    def snp_planes(self, Z: Tensor, row_index: Tensor, d: engine.DeltaResult | None = None, *,
                   keep_us: bool = False):
        """Scan planes S1 = B^T s, S2 = (B^T s)^2 of the SNP columns Z[:, row_index] (core.py:628-631), or
        S2 = (B^T s)^2 A for the factored weights of d (default: the last fit); keep_us (FP64 SNP rotation only)
        also returns the FP64 rotated SNPs US (rows_pad, kpad)."""
        d = self.delta if d is None else d
        fac = None if d is None else d.fac
        mc = row_index.numel()
        rot = engine.settings.snp_rotation
        if rot == "fp64":
            us = engine.rotate_snps_f64(Z, row_index, mc, self.mat, self.col0, self.neff, rup(mc, TILE), self.kpad)
            s1, s2 = engine.snp_planes_from_f64(self.prec, us, mc, self.neff, fac)
            return (s1, s2, us) if keep_us else (s1, s2)
        if keep_us:
            raise ValueError("keep_us needs the FP64 SNP rotation")
        if rot != "split":
            raise ValueError(f"unknown snp_rotation setting {rot!r}")
        spl = engine.pack(self.prec, Z, mc, self.N, 1, Z.shape[1], rup(mc, TILE), self.npad, row_index=row_index)
        if fac is not None:
            return engine.factored_snp_planes(spl, self.bt, mc, fac)
        return engine.rotate_us(spl, self.bt)


def _projector(X: Tensor) -> tuple[Tensor, Tensor]:
    """X^+ (torch.linalg.pinv, core.py:498) and I - X X^+ (core.py:499-500)."""
    N = X.shape[0]
    Xpinv = torch.linalg.pinv(X)
    return Xpinv, torch.eye(N, dtype=X.dtype, device=X.device) - X @ Xpinv


_SELECT_CAP = [1 << 16]  # list length of the last selection (grown on demand)


def _select_above(F: Tensor, thr: float) -> tuple[Tensor, Tensor]:
    """(row, column) int64 pairs of the entries of F (rows x cols FP64, unit column stride) above thr, in an
    arbitrary order (flmm_select_above; one host read of the count)."""
    rows, cols = F.shape
    cnt = torch.zeros(1, dtype=torch.int64, device=F.device)
    while True:
        cap = _SELECT_CAP[0]
        mi = torch.empty(cap, dtype=torch.int64, device=F.device)
        pi = torch.empty(cap, dtype=torch.int64, device=F.device)
        _lib.call("flmm_select_above", F.data_ptr(), rows, cols, F.stride(0), float(thr), mi.data_ptr(),
                  pi.data_ptr(), cap, cnt.data_ptr(), _lib.stream())
        n = int(cnt.item())
        if n <= cap:
            return mi[:n], pi[:n]
        _SELECT_CAP[0] = rup(n, 1 << 16)
        cnt.zero_()


def _project(K: Tensor, X: Tensor, Xpinv: Tensor, Iperp: Tensor) -> Tensor:
    """Iperp (K + I) Iperp (core.py:503-510) for symmetric K, with H = X X^+ = I - Iperp as rank-C updates:
    K - H K - K H + H K H + Iperp, H K = X (X^+ K), H K H = X (X^+ K X) X^+ (O(N^2 C) instead of two N^3
    products; equal to FP64 rounding)."""
    G = Xpinv @ K                 # C x N
    T = X @ G                     # H K
    M = (G @ X) @ Xpinv           # C x N: X^+ K X X^+
    return K - T - T.T + X @ M + Iperp


def _projected_eigh(K: Tensor, Iperp: Tensor, D: int, shift: float = 1.0) -> tuple[Tensor, Tensor]:
    """eigh of Iperp (K + shift I) Iperp (core.py:503-510); returns (s = clamp(lambda[D:] - shift, 0), U)."""
    N = K.shape[0]
    Kp = (Iperp @ (K + shift * torch.eye(N, dtype=K.dtype, device=K.device))) @ Iperp
    s_full, U = _eigh(Kp)
    return torch.clamp(s_full[D:N] - shift, min=0.0), U


def _spec_state(spec) -> _Basis:
    st = getattr(spec, "_gpu", None)
    if st is None:
        raise ValueError("spectrum was not built by fasterlmm_amd.core (rotate / fastlmm_compat_rotate)")
    return st


def _kinship(Z: Tensor) -> Tensor:
    """K = Z Z^T / M of a standardised panel (io.grm's product, io.py:257-265), the Gram by the kinship setting
    (engine.gram: the FP64 Gram kernel by default)."""
    return engine.gram(Z) / float(Z.shape[1])


def _new_outputs(M, n_real, P, dev, *, zero: bool = True):
    return engine.ScanOutputs.allocate(M, n_real, P, dev, zero=zero)


# COMPAT PATH -------

def fastlmm_compat_rotate(K: Tensor, X: Tensor, Y: Tensor) -> CompatSpectrum:
    """Projected rotation (core.py:487-519): eigh of (I - X X^+)(K + I)(I - X X^+), drop the first D = ncols(X)
    eigenpairs, s = clamp(lambda - 1, 0), UY = U_eff^T (Y - X X^+ Y) (the product on MFMA)."""
    home, (Kd, Xd, Yd) = _dev(K, X, Y)
    N = Kd.shape[0]
    D = Xd.shape[1]
    Xpinv, Iperp = _projector(Xd)
    s, U = _projected_eigh(Kd, Iperp, D)
    st = _Basis(s, U, D, N - D, engine.prec_code(), nu=N - D - 1)
    st.rotate_columns(Yd - Xd @ (Xpinv @ Yd), in_span=True)
    spec = CompatSpectrum(s=s.to(home), U_eff=U[:, D:N].contiguous().to(home), X=Xd.to(home), Xpinv=Xpinv.to(home),
                          UY=st.uy_dense().to(home))
    spec._gpu = st
    return spec


def fit_delta_grid_compat(spec: CompatSpectrum, *, n_grid: int = 64, log_delta_min: float = -10.0,
                          log_delta_max: float = 10.0, refine: bool = True) -> Tensor:
    """Per-column log delta for the compat path: grid argmin then refinement in [g[i-1], g[i+1]]
    (core.py:561-595).  Returns (P,) FP64 on the spectrum's device."""
    st = _spec_state(spec)
    ld, _ = st.fit(n_grid, float(log_delta_min), float(log_delta_max), refine)
    return ld.to(spec.s.device)


def snp_wald_scan_compat(spec: CompatSpectrum, log_delta: Tensor, S: Tensor, *, snp_chunk: int = 4096,
                         pheno_chunk: int = 256, n_real: int | None = None) -> ScanResult:
    """Wald F / beta / SE / SnpFractVarExpl of every SNP column of S against every column of the spectrum, in the
    projected basis with dof Neff - 1 (core.py:601-671).  The first n_real columns keep full detail, every
    column keeps its maximum F."""
    st = _spec_state(spec)
    home, (Sd, ldd) = _dev(S, log_delta)
    P = st.P
    M = Sd.shape[1]
    n_real = P if n_real is None else min(int(n_real), P)
    d = st.delta_for(ldd, n_real)
    out = _new_outputs(M, n_real, P, Sd.device)
    st.scan(Sd.contiguous(), torch.arange(M, device=Sd.device, dtype=torch.int64), n_real, out, d)
    engine.refresh_real_max(out, n_real)
    return _home(ScanResult(out.f, out.beta, out.se, out.sfve, out.nullh2, out.max_F), home)


def _loco(Z, X, Y, chrom, n_real, on_chrom, Z_kin, chrom_kin, n_grid, *, plain: bool, reml: bool):
    """Shared LOCO driver (core.py:403-436, 694-726): per chromosome the kinship of every other chromosome
    (of Z, or of the kinship panel), the projected basis, per-column delta and the scan of the chromosome's
    SNPs; real-column rows are scattered into place and max_F is the maximum over chromosomes."""
    if (Z_kin is None) != (chrom_kin is None):
        raise ValueError("Z_kin and chrom_kin go together")
    home, (Zd, Xd, Yd, Zkd) = _dev(Z, X, Y, Z_kin)
    chrom = np.asarray(chrom)
    N, M = Zd.shape
    P = Yd.shape[1]
    D = Xd.shape[1]
    n_real = P if n_real is None else min(int(n_real), P)
    dev = Zd.device
    prec = engine.prec_code()
    out = _new_outputs(M, n_real, P, dev, zero=False)  # every SNP belongs to one chromosome: every row is written
    Zk = Zd if Zkd is None else Zkd
    ck = chrom if chrom_kin is None else np.asarray(chrom_kin)
    stage = engine.stage
    chroms = sorted(np.unique(chrom).tolist())
    # every chromosome's SNP rows in one upload before any kernel is queued (a copy from host memory waits for the
    # stream), then device slices
    rows_of = [np.flatnonzero(chrom == c) for c in chroms]
    all_rows = torch.from_numpy(np.concatenate(rows_of) if rows_of else np.zeros(0, np.int64)).to(dev)
    row_start = np.concatenate([[0], np.cumsum([r.size for r in rows_of])])
    with stage("kinship"):
        # each kinship chromosome's Gram once; K_all = their sum, in chromosome order (the Gram of all markers to
        # FP64 rounding), and K_loco(c) = (K_all - G_c) / (M - m_c) below
        grams = {}
        for c in sorted(np.unique(ck).tolist()):
            cols = np.flatnonzero(ck == c)
            if cols[-1] - cols[0] + 1 == cols.size:  # consecutive columns: a view, no gather
                Zc = Zk[:, int(cols[0]):int(cols[-1]) + 1]
            else:
                Zc = Zk.index_select(1, torch.from_numpy(cols).to(dev))
            grams[c] = (engine.gram(Zc, prec), cols.size)
            del Zc
        K_all = None
        for G, _ in grams.values():
            K_all = G.clone() if K_all is None else K_all.add_(G)
    Xpinv, Iperp = _projector(Xd)
    # the basis is orthogonal to X, so B^T (Y - X X^+ Y) = B^T Y; the OLS regress-out keeps the compat
    # path's arithmetic (core.py:517) and costs one pass
    with stage("pack_columns"):
        Bc = Xpinv @ Yd
        if D <= 8:  # Y_r = Y - X Bc is packed in one pass, never formed
            ypl = engine.pack_columns_projected(prec, Yd, Xd, Bc, rup(max(P, 1), PTILE), rup(N, KCHUNK))
            y_real = torch.addmm(Yd[:, :n_real], Xd, Bc[:, :n_real], alpha=-1.0)  # kept for the FP64 re-scoring
        else:
            Y_r = torch.addmm(Yd, Xd, Bc, alpha=-1.0)
            ypl = engine.pack(prec, Y_r, P, N, 1, P, rup(max(P, 1), PTILE), rup(N, KCHUNK))
            y_real = Y_r[:, :n_real].contiguous()
            del Y_r
        del Bc
    jitter = 1e-6 if plain else 0.0
    eye = torch.eye(N, dtype=torch.float64, device=dev)
    # every chromosome's projected kinship first, then one batched FP64 eigendecomposition
    Kps, Ks = [], []
    with stage("kinship"):
        for c in chroms:
            G, m_c = grams.pop(c, (None, 0))
            K = ((K_all - G) if G is not None else K_all) / float(Zk.shape[1] - m_c)
            del G
            if jitter:
                K = K + jitter * eye
            Kps.append(_project(K, Xd, Xpinv, Iperp))
            if plain and not reml:
                Ks.append(K)
        del grams, K_all
    with stage("eigh"):
        ev, Ub = _eigh(torch.stack(Kps))
        del Kps
        s_mls = [torch.clamp(x, min=0.0) for x in torch.linalg.eigvalsh(torch.stack(Ks))] if Ks else None
        del Ks

    def prepare(k: int):
        """Chromosome k's basis, rotated columns and delta fit, on the current stream."""
        s = torch.clamp(ev[k, D:N] - 1.0, min=0.0)
        with stage("pack_basis"):
            st = _Basis(s, Ub[k], D, N - D, prec, nu=N - D - 1, s_ml=None if s_mls is None else s_mls[k])
        with stage("rotate_columns"):
            st.use_column_planes(ypl, P, y_real, in_span=True)
        with stage("delta"):
            _, d = st.fit(n_grid, -10.0, 10.0, True, projected_loss=(not plain) or reml, n_real=n_real)
        done = torch.cuda.Event()
        done.record()
        return st, d, done

    # With engine.settings.pipeline, chromosome k + 1 is prepared on a second stream while chromosome k scans (the
    # delta search and rotations beside the MFMA-bound scan); each chromosome's buffers are then dropped only after
    # the scan stream has been synchronised, so the caching allocator never hands out memory a queued kernel on the
    # other stream still reads.  Otherwise everything runs on the current stream, in order, where freed memory is
    # only reused by later work on that stream, and the host does not wait between chromosomes.
    main = torch.cuda.current_stream(dev)
    side = torch.cuda.Stream(device=dev) if engine.settings.pipeline else main
    side.wait_stream(main)
    with torch.cuda.stream(side):
        nxt = prepare(0)
    for k, c in enumerate(chroms):
        st, d, done = nxt
        nxt = None
        main.wait_event(done)
        rows = rows_of[k]
        test = all_rows[int(row_start[k]):int(row_start[k + 1])]
        first = int(rows[0]) if rows.size and rows[-1] - rows[0] + 1 == rows.size else None  # consecutive rows

        def launch_next(k=k):
            nonlocal nxt
            if k + 1 < len(chroms):
                with torch.cuda.stream(side):
                    nxt = prepare(k + 1)

        piped = side is not main  # serially, the next chromosome is prepared after this scan, outside its stage
        with stage("snps+scan"):
            st.scan(Zd, test, n_real, out, d, after_launch=launch_next if piped else None, first_row=first)
        if not piped:
            launch_next()
        if piped or on_chrom is not None:
            main.synchronize()
        del st, d, done
        if on_chrom is not None:
            on_chrom(k + 1, len(chroms))
    engine.refresh_real_max(out, n_real)
    return _home(ScanResult(out.f, out.beta, out.se, out.sfve, out.nullh2, out.max_F), home)


def _profile_loss_compat(log_delta: Tensor, s: Tensor, UY: Tensor) -> Tensor:
    """(G, P) compat -2 log-likelihood Neff (log 2 pi + log(max(sum_k UY^2 / (s_k + delta), floor) / Neff) + 1) +
    sum_k log(s_k + delta) on a grid of log delta (core.py:525-542), FP64 on the inputs' device.  The scan's own
    delta search evaluates this loss inside the delta kernels; this is the reference's helper for its callers."""
    neff = s.shape[0]
    Sd = s.unsqueeze(0) + log_delta.exp().unsqueeze(1)
    rWr = torch.clamp((1.0 / Sd) @ (UY * UY), min=_pos_floor(s.dtype))
    log2pi = math.log(2.0 * math.pi)
    return neff * (log2pi + (rWr / neff).log() + 1.0) + Sd.log().sum(dim=1).unsqueeze(1)


def _profile_loss_compat_per_pheno(log_delta: Tensor, s: Tensor, UY: Tensor) -> Tensor:
    """(P,) compat loss with one log delta per column (core.py:545-558)."""
    neff = s.shape[0]
    Sd = s.unsqueeze(0) + log_delta.exp().unsqueeze(1)
    rWr = torch.clamp(((1.0 / Sd) * UY.T * UY.T).sum(dim=1), min=_pos_floor(s.dtype))
    log2pi = math.log(2.0 * math.pi)
    return neff * (log2pi + (rWr / neff).log() + 1.0) + Sd.log().sum(dim=1)


def loco_scan_compat(Z: Tensor, X: Tensor, Y: Tensor, chrom: np.ndarray, *, n_real: int | None = None,
                     on_chrom=None, Z_kin: Tensor | None = None, chrom_kin: np.ndarray | None = None,
                     n_grid: int = 64) -> ScanResult:
    """Leave-one-chromosome-out compat scan (core.py:677-726).  Z must be pre-standardised; Z_kin / chrom_kin
    (both or neither) build every K_loco from a separate pre-standardised panel.  n_grid defaults to the
    reference's 64 (fit_delta_grid_compat's default)."""
    return _loco(Z, X, Y, chrom, n_real, on_chrom, Z_kin, chrom_kin, n_grid, plain=False, reml=False)


def single_k_scan_compat(Z: Tensor, X: Tensor, Y: Tensor, *, n_real: int | None = None,
                         Z_kin: Tensor | None = None, n_grid: int = 64) -> ScanResult:
    """One kinship over all SNPs (or the Z_kin panel), compat rotation, delta fit, scan of every SNP
    (core.py:729-740)."""
    home, (Zd, Zkd) = _dev(Z, Z_kin)
    K = _kinship(Zd if Zkd is None else Zkd)
    spec = fastlmm_compat_rotate(K, X, Y)
    ld = fit_delta_grid_compat(spec, n_grid=n_grid)
    return _home(snp_wald_scan_compat(spec, ld, Zd, n_real=n_real), home)


# PLAIN GLS PATH -------

def _require_full_rank(X, what: str = "X") -> None:
    """ValueError unless the covariate design has full column rank.  The plain GLS path's Cholesky of
    X~^T V^-1 X~ (core.py:166-171, 203-207, 336-340) fails on a singular Gram at some deltas and not at others,
    and the reference then substitutes an identity factor (scan) and an infinite loss (delta fit): its results
    for such an X depend on rounding, so this build refuses them instead of reproducing them."""
    Xt = torch.as_tensor(X).to(torch.float64)
    C = Xt.shape[1]
    r = int(torch.linalg.matrix_rank(Xt)) if C > 0 and Xt.shape[0] > 0 else C
    if r < C:
        raise ValueError(f"the plain GLS path needs a full-column-rank covariate design {what}: rank {r} < {C} "
                         f"columns (the compat path, loco_scan_compat / single_k_scan_compat, handles it)")


def _plain_basis(s: Tensor, X_rot: Tensor, prec: int) -> tuple[_Basis, Tensor]:
    """Projected basis inside the rotated space of a Spectrum: Q = orthonormal basis of null(X_rot^T) (complete
    QR), eigh of Q^T diag(s) Q, B = Q V.  Returns (basis state, B (N, N - C))."""
    N, C = X_rot.shape
    Q, _ = torch.linalg.qr(X_rot, mode="complete")
    Qn = Q[:, C:]
    Bm = Qn.T @ (s.unsqueeze(1) * Qn)
    sp, V = _eigh(0.5 * (Bm + Bm.T))
    B = (Qn @ V).contiguous()
    st = _Basis(torch.clamp(sp, min=0.0), B, 0, N - C, prec, nu=N - C - 1, s_ml=s)
    return st, B


def rotate(K: Tensor, X: Tensor, Y: Tensor) -> Spectrum:
    """Spectrum: s, U = eigendecompose(K) (K + 1e-6 I, clamped), X_rot = U^T X, Y_rot = U^T Y (core.py:105-116)."""
    _require_full_rank(X)
    home, (Kd, Xd, Yd) = _dev(K, X, Y)
    s, U = eigendecompose(Kd)
    X_rot = U.T @ Xd
    st, _ = _plain_basis(s, X_rot, engine.prec_code())
    # rotated-space columns of the projected basis: B^T (U^T Y) = (U B)^T Y
    Y_rot = U.T @ Yd
    st.rotate_columns(Y_rot)
    spec = Spectrum(s=s.to(home), U=U.to(home), X_rot=X_rot.to(home), Y_rot=Y_rot.to(home))
    spec._gpu = st
    return spec


def _spectrum_state(spectrum: Spectrum) -> _Basis:
    st = getattr(spectrum, "_gpu", None)
    if st is None:
        _require_full_rank(spectrum.X_rot, "X_rot")
        _, (s, X_rot, Y_rot) = _dev(spectrum.s, spectrum.X_rot, spectrum.Y_rot)
        st, _ = _plain_basis(s, X_rot, engine.prec_code())
        st.rotate_columns(Y_rot)
        spectrum._gpu = st
    return st


def nLLeval(delta, s: Tensor, X_rot: Tensor, y_rot: Tensor) -> Tensor:
    """Profiled ML -2 log-likelihood of one column at one delta (core.py:122-145), FP64."""
    N, _ = X_rot.shape
    Sd = s + delta
    w = 1.0 / Sd
    WX = w.unsqueeze(-1) * X_rot
    A = WX.T @ X_rot
    Xy = WX.T @ y_rot
    beta = torch.linalg.solve(A, Xy)
    rWr = (y_rot * w * y_rot).sum() - (Xy * beta).sum()
    log2pi = torch.log(torch.tensor(2.0 * torch.pi, dtype=s.dtype, device=s.device))
    return N * (log2pi + (rWr / N).log() + 1.0) + Sd.log().sum()


def _profile_loss(log_delta: Tensor, s: Tensor, X_rot: Tensor, Y_rot: Tensor, *, reml: bool = False) -> Tensor:
    """(G, P) ML / REML -2 log-likelihood on a grid (core.py:148-187), FP64 on the input's device."""
    N, C = X_rot.shape
    delta = log_delta.exp()
    Sd = s.unsqueeze(0) + delta.unsqueeze(1)
    w = 1.0 / Sd
    WX = w.unsqueeze(2) * X_rot.unsqueeze(0)
    A = torch.einsum("gnc,nd->gcd", WX, X_rot)
    L, info = torch.linalg.cholesky_ex(A)
    bad = info > 0
    if bad.any():
        L = torch.where(bad.view(-1, 1, 1), torch.eye(C, dtype=A.dtype, device=A.device).expand_as(L), L)
    logdet_A = 2.0 * torch.diagonal(L, dim1=-2, dim2=-1).log().sum(-1)
    u = torch.einsum("gnc,np->gcp", WX, Y_rot)
    uAu = (u * torch.cholesky_solve(u, L)).sum(dim=1)
    rWr = torch.clamp(w @ (Y_rot * Y_rot) - uAu, min=_pos_floor(s.dtype))
    sum_log_Sd = Sd.log().sum(dim=1)
    log2pi = torch.log(torch.tensor(2.0 * torch.pi, dtype=s.dtype, device=s.device))
    if reml:
        loss = (N - C) * (log2pi + (rWr / (N - C)).log() + 1.0) + sum_log_Sd.unsqueeze(1) + logdet_A.unsqueeze(1)
    else:
        loss = N * (log2pi + (rWr / N).log() + 1.0) + sum_log_Sd.unsqueeze(1)
    if bad.any():
        loss = torch.where(bad.unsqueeze(1), torch.full_like(loss, float("inf")), loss)
    return loss


def _profile_loss_per_pheno(log_delta: Tensor, s: Tensor, X_rot: Tensor, Y_rot: Tensor, *,
                            reml: bool = False) -> Tensor:
    """(P,) loss with one log delta per column (core.py:190-224)."""
    G = torch.stack([_profile_loss(log_delta[p:p + 1], s, X_rot, Y_rot[:, p:p + 1], reml=reml)[0, 0]
                     for p in range(Y_rot.shape[1])])
    return G


def fit_delta_grid(spectrum: Spectrum, *, n_grid: int = 64, log_delta_min: float = -10.0,
                   log_delta_max: float = 10.0, refine: bool = True, reml: bool = False) -> Tensor:
    """Per-column log delta for the GLS path, ML or REML: grid argmin then refinement in [g[i-1], g[i+1]]
    (core.py:227-264), on the projected basis (ML keeps the full log det V of the spectrum and N)."""
    st = _spectrum_state(spectrum)
    ld, _ = st.fit(n_grid, float(log_delta_min), float(log_delta_max), refine, projected_loss=reml)
    return ld.to(spectrum.s.device)


def _auto_tile(N: int, C: int, M: int, P: int, device: torch.device) -> tuple[int, int]:
    """The reference's tile heuristic (core.py:270-292), kept for API compatibility; the kernels tile
    internally in 128 x 128 MFMA tiles and ignore it."""
    if device.type != "cuda":
        return min(256, P), min(4096, M)
    free_bytes = torch.cuda.mem_get_info(device)[0]
    budget = int(free_bytes * 0.5)
    T = 12
    snp_chunk = min(M, 16384)
    while snp_chunk >= 512:
        per_pc = 8 * (N * C + (2 * C + T) * snp_chunk)
        pheno_chunk = (budget - 8 * N * snp_chunk) // max(per_pc, 1)
        if pheno_chunk >= 64:
            return int(min(pheno_chunk, P)), int(snp_chunk)
        snp_chunk //= 2
    return min(64, P), min(512, M)


def snp_wald_scan(spectrum: Spectrum, log_delta: Tensor, S_rot: Tensor, *, snp_chunk: int | None = None,
                  pheno_chunk: int | None = None, n_real: int | None = None) -> ScanResult:
    """GLS Wald F / beta / SE / SnpFractVarExpl of pre-rotated SNPs S_rot = U^T Z against every column, dof
    N - C - 1 (core.py:295-375), through the projected basis of the spectrum."""
    st = _spectrum_state(spectrum)
    home, (Sd, ldd) = _dev(S_rot, log_delta)
    P = st.P
    M = Sd.shape[1]
    n_real = P if n_real is None else min(int(n_real), P)
    d = st.delta_for(ldd, n_real)
    out = _new_outputs(M, n_real, P, Sd.device)
    st.scan(Sd.contiguous(), torch.arange(M, device=Sd.device, dtype=torch.int64), n_real, out, d)
    engine.refresh_real_max(out, n_real)
    return _home(ScanResult(out.f, out.beta, out.se, out.sfve, out.nullh2, out.max_F), home)


def loco_scan(Z: Tensor, X: Tensor, Y: Tensor, chrom: np.ndarray, *, n_real: int | None = None, on_chrom=None,
              Z_kin: Tensor | None = None, chrom_kin: np.ndarray | None = None, n_grid: int = 64,
              reml: bool = False) -> ScanResult:
    """Leave-one-chromosome-out GLS scan (core.py:381-436): per chromosome K_loco + 1e-6 I, ML (or REML) delta
    per column and the GLS Wald test of the chromosome's SNPs.  X must have full column rank (ValueError)."""
    _require_full_rank(X)
    return _loco(Z, X, Y, chrom, n_real, on_chrom, Z_kin, chrom_kin, n_grid, plain=True, reml=reml)


def single_k_scan(Z: Tensor, X: Tensor, Y: Tensor, *, n_real: int | None = None,
                  Z_kin: Tensor | None = None) -> ScanResult:
    """One kinship, GLS rotation, delta fit and scan of every SNP (core.py:439-451)."""
    _require_full_rank(X)
    home, (Zd, Zkd) = _dev(Z, Z_kin)
    spec = rotate(_kinship(Zd if Zkd is None else Zkd), X, Y)
    ld = fit_delta_grid(spec)
    S_rot = _dev(spec.U)[1][0].T @ Zd
    return _home(snp_wald_scan(spec, ld, S_rot, n_real=n_real), home)
