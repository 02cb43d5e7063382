"""GPU pipeline of the LMM scan: operand planes, rotations, delta search and the fused Wald scan.

Every array here is a torch tensor on the GPU used as device memory; all arithmetic of the hot path runs in the
hand-written kernels of libfasterlmm_amd.so (called through ``_lib``).  torch is used for allocation, for the
small FP64 setup products (covariate projector, N x N projection) that rocBLAS runs, and for
rocSOLVER's FP64 symmetric eigendecomposition.

Data layout (HBM)
  genotypes     Z (N, M) FP64 row-major, standardised (the reference's input layout)
  planes        operand matrices in K-chunks of 32 (see include/fasterlmm_amd.h): Y_r columns (P_pad x N_pad),
                U_eff^T (K_pad x N_pad), rotated columns UY (fp32, P_pad x K_pad), SNP planes S1 = US, S2 = US^2
                (Mc_pad x K_pad) and column planes Q1 = w o UY, Q2 = w (P_pad x K_pad); with factored weights
                (expsum.py) S2 = US^2 A (Mc_pad x R_pad) and Q2 = B (P_pad x R_pad), R_pad = 128 typically
  outputs       f, beta, se, sfve, nullh2 (M, n_real) FP64 row-major; max_F (P,) FP64
with N_pad = roundup(N, 32), K_pad = roundup(Neff, 128), P_pad = roundup(P, 128), Mc_pad = roundup(Mc, 128).
"""
from __future__ import annotations

import contextlib
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, expsum

TILE = 128
PTILE = 256  # column padding: the wide scan variant tiles columns by 256
KCHUNK = 32


def rup(x: int, m: int) -> int:
    return (int(x) + m - 1) // m * m


@dataclass
class Settings:
    precision: str = "split_fp16"  # "split_fp16" (scaled FP16 hi+lo, 3 MFMA products), "split_bf16" (BF16 hi+lo,
                                   # 3 products) or "fp32" (FP32 MFMA)
    delta: str = "newton"          # "newton" (exact local minimiser), "golden" (the reference's 50 steps) or
                                   # "learned" (MLP proposal + checked Newton, grid as fallback and arbiter)
    weights: str = "auto"          # sWs weights 1 / (s + delta): "factored" (exponential sum, second product over
                                   # ~128 terms), "exact" (over Neff), "auto" = factored for the split precisions
    kinship: str = "auto"          # kinship Gram Z Z^T: "fp64" (FP64 MFMA kernel), "blas" (rocBLAS FP64), "split"
                                   # (MFMA in the split precision, FP64 folds every 1,024 k; ~1e-7 of the diagonal,
                                   # which moves beta by ~1e-4 where delta sits near 1e-4 next to eigenvalues of
                                   # that size), "auto" = fp64
    snp_rotation: str = "fp64"     # SNP side US = U_eff^T S: "fp64" (FP64 MFMA kernel, then packed) or "split" (one
                                   # fused MFMA pass in the split precision: ~2^-22 of |S| in every direction, which
                                   # the scan weighs by 1 / (s_k + delta))
    grid: str = "auto"             # delta grid GEMM: "auto" = split-FP16 MFMA where the rotated columns carry plane
                                   # scales (split_fp16), "fp32" = FP32 MFMA
    screen: str = "auto"           # permutation columns' max_F: "auto" = the certified hi-plane screen with FP64
                                   # re-evaluation of the tests that can hold each maximum (split_fp16 with the
                                   # FP64 SNP rotation), every column on the sWy-only tile with 1 / sWs
                                   # interpolated in its log-delta bucket; "general" = the same screen with both
                                   # products on hi planes; "off" = the full split scan for every column
    pipeline: bool = False         # LOCO: prepare chromosome c + 1 (basis, rotation, delta) on a second stream
                                   # while chromosome c scans (same bits; ~1% less step time on the N=1000,
                                   # M=100k workload, but the co-running kernels slow the scan kernel ~1.7x)


settings = Settings()


@contextlib.contextmanager
def configure(**kw):
    """Temporarily change ``settings`` (precision=..., delta=...)."""
    old = {k: getattr(settings, k) for k in kw}
    for k, v in kw.items():
        if not hasattr(settings, k):
            raise ValueError(f"unknown setting {k!r}")
        setattr(settings, k, v)
    try:
        yield settings
    finally:
        for k, v in old.items():
            setattr(settings, k, v)


def prec_code(precision: str | None = None) -> int:
    p = settings.precision if precision is None else precision
    if p == "split_fp16":
        return _lib.PREC_SPLIT_FP16
    if p == "split_bf16":
        return _lib.PREC_SPLIT_BF16
    if p == "fp32":
        return _lib.PREC_FP32
    raise ValueError(f"unknown precision {p!r}")


def delta_code(mode: str | None = None) -> int:
    m = settings.delta if mode is None else mode
    if m == "newton":
        return _lib.DELTA_NEWTON
    if m == "golden":
        return _lib.DELTA_GOLDEN
    if m == "learned":  # the learned path's fallback / fixed refinements are Newton in the grid bracket
        return _lib.DELTA_NEWTON
    raise ValueError(f"unknown delta mode {m!r}")


class Planes:
    """One operand matrix as MFMA planes: a (hi, lo) bf16 or fp16 pair (fp16 rows carry a power-of-two scale)
    or a single fp32 plane."""

    def __init__(self, prec: int, rows_pad: int, kpad: int, dev, *, fp32_only: bool = False):
        self.prec = _lib.PREC_FP32 if fp32_only else prec
        self.rows_pad, self.kpad = rows_pad, kpad
        n = rows_pad * kpad
        self.scale = None
        if self.prec in (_lib.PREC_SPLIT_BF16, _lib.PREC_SPLIT_FP16):
            self.hi = torch.empty(n, dtype=torch.int16, device=dev)
            self.lo = torch.empty(n, dtype=torch.int16, device=dev)
            if self.prec == _lib.PREC_SPLIT_FP16:
                self.scale = torch.empty(rows_pad, dtype=torch.float32, device=dev)
        else:
            self.hi = torch.empty(n, dtype=torch.float32, device=dev)
            self.lo = None

    @property
    def ptrs(self):
        return self.hi.data_ptr(), (self.lo.data_ptr() if self.lo is not None else None)

    @property
    def scale_ptr(self):
        return None if self.scale is None else self.scale.data_ptr()

    def to_dense(self, rows: int, k: int) -> torch.Tensor:
        """(rows, k) FP64 view of the stored values (hi + lo, unscaled, for split planes)."""
        def one(t, dt):
            v = t.view(dt) if dt is not None else t
            return v.reshape(self.kpad // KCHUNK, self.rows_pad, KCHUNK).permute(1, 0, 2).reshape(
                self.rows_pad, self.kpad)[:rows, :k]
        if self.prec == _lib.PREC_FP32:
            return one(self.hi, None).double()
        dt = torch.float16 if self.prec == _lib.PREC_SPLIT_FP16 else torch.bfloat16
        x = one(self.hi, dt).double() + one(self.lo, dt).double()
        if self.scale is not None:
            x = x / self.scale[:rows].double().unsqueeze(1)
        return x


def fp32_plane_to_dense(plane: torch.Tensor, rows_pad: int, kpad: int, rows: int, k: int) -> torch.Tensor:
    return plane.reshape(kpad // KCHUNK, rows_pad, KCHUNK).permute(1, 0, 2).reshape(rows_pad, kpad)[:rows, :k]


def pack(prec: int, src: torch.Tensor, n_rows: int, n_k: int, row_stride: int, k_stride: int, rows_pad: int,
         kpad: int, *, row_index: torch.Tensor | None = None, offset: int = 0) -> Planes:
    """Planes of element (r, k) = src.flat[offset + src_row(r) * row_stride + k * k_stride]."""
    assert src.dtype == torch.float64 and src.is_cuda
    pl = Planes(prec, rows_pad, kpad, src.device)
    base = src.data_ptr() + 8 * offset
    hi, lo = pl.ptrs
    norm = None
    st = _lib.stream()
    if pl.prec == _lib.PREC_SPLIT_FP16:
        norm = torch.empty(max(n_rows, 1), dtype=torch.float64, device=src.device)
        _lib.call("flmm_row_norms", base, n_rows, n_k, row_stride, k_stride, _lib.ptr(row_index), norm.data_ptr(), st)
    _lib.call("flmm_pack_planes", pl.prec, base, n_rows, n_k, row_stride, k_stride, _lib.ptr(row_index),
              _lib.ptr(norm), hi, lo, pl.scale_ptr, rows_pad, kpad, st)
    return pl


def pack_columns_projected(prec: int, Y: torch.Tensor, X: torch.Tensor, B: torch.Tensor, rows_pad: int,
                           kpad: int) -> Planes:
    """Planes of the transposed residual Y - X B (row p = column p, k = strain) without forming it
    (flmm_pack_columns_projected): Y (N, P), X (N, C), B (C, P) FP64 on the GPU."""
    Y, X, B = (t.contiguous() for t in (Y, X, B))
    N, P = Y.shape
    pl = Planes(prec, rows_pad, kpad, Y.device)
    hi, lo = pl.ptrs
    norm = torch.empty(max(P, 1), dtype=torch.float64, device=Y.device) if prec == _lib.PREC_SPLIT_FP16 else None
    _lib.call("flmm_pack_columns_projected", prec, Y.data_ptr(), N, P, X.data_ptr(), X.shape[1], B.data_ptr(),
              _lib.ptr(norm), hi, lo, pl.scale_ptr, rows_pad, kpad, _lib.stream())
    return pl


def gram(Z: torch.Tensor, prec: int | None = None) -> torch.Tensor:
    """G = Z Z^T (n x n FP64) of Z (n, m) FP64 on the GPU, by the kinship setting: flmm_gram_f64 on the FP64
    matrix cores (default), rocBLAS ("blas"), or flmm_gram in the split precision ("split").  Deterministic: the same
    Z gives the same bits."""
    prec = prec_code() if prec is None else prec
    k = settings.kinship
    if k not in ("auto", "fp64", "blas", "split"):
        raise ValueError(f"unknown kinship setting {k!r}")
    if k == "blas":
        return Z @ Z.T
    n, m = Z.shape
    if m == 0 or n == 0:
        return torch.zeros(n, n, dtype=torch.float64, device=Z.device)
    if Z.stride(1) != 1:
        Z = Z.contiguous()
    ldz = Z.stride(0)  # a column range of a wider panel is read in place
    rows_pad = rup(n, TILE)
    t = rows_pad // TILE
    pairs = t * (t + 1) // 2
    if k in ("auto", "fp64"):
        n_split = max(1, min(-(-1024 // pairs), -(-m // 256)))  # >= 1,024 workgroups where K allows
        partial = torch.empty(n_split, rows_pad, rows_pad, dtype=torch.float64, device=Z.device)
        out = torch.empty(n, n, dtype=torch.float64, device=Z.device)
        _lib.call("flmm_gram_f64", Z.data_ptr(), ldz, n, m, n_split, partial.data_ptr(), out.data_ptr(), n,
                  _lib.stream())
        return out
    if prec not in (_lib.PREC_SPLIT_FP16, _lib.PREC_SPLIT_BF16):
        raise ValueError("the split kinship needs a split precision")
    kpad = rup(m, KCHUNK)
    pl = pack(prec, Z, n, m, ldz, 1, rows_pad, kpad)
    n_split = max(1, min(-(-1024 // pairs), -(-(kpad // KCHUNK) // 32)))  # >= 1,024 workgroups where K allows
    partial = torch.empty(n_split, rows_pad, rows_pad, dtype=torch.float64, device=Z.device)
    out = torch.empty(n, n, dtype=torch.float64, device=Z.device)
    hi, lo = pl.ptrs
    _lib.call("flmm_gram", pl.prec, hi, lo, pl.scale_ptr, rows_pad, kpad, n, n_split, partial.data_ptr(),
              out.data_ptr(), n, _lib.stream())
    return out


def rotate_us_x2(a: Planes, b: Planes) -> tuple[Planes, torch.Tensor]:
    """S1 = x planes and x^2 as FP64 (a.rows_pad, b.rows_pad) row-major, x[i][j] = sum_n a[i][n] b[j][n]."""
    dev = a.hi.device
    s1 = Planes(a.prec, a.rows_pad, b.rows_pad, dev)
    x2 = torch.empty(a.rows_pad, b.rows_pad, dtype=torch.float64, device=dev)
    ah, al = a.ptrs
    bh, bl = b.ptrs
    o1 = s1.lo.data_ptr() if s1.lo is not None else None
    _lib.call("flmm_rotate", a.prec, _lib.ROTATE_US_X2, ah, al, a.scale_ptr, a.rows_pad, bh, bl, b.scale_ptr,
              b.rows_pad, a.kpad, s1.hi.data_ptr(), o1, x2.data_ptr(), None, s1.scale_ptr, None, _lib.stream())
    return s1, x2


def rotate_uy(a: Planes, b: Planes) -> torch.Tensor:
    """fp32 plane (a.rows_pad x b.rows_pad) of out[i][j] = sum_n a[i][n] b[j][n]."""
    out = torch.empty(a.rows_pad * b.rows_pad, dtype=torch.float32, device=a.hi.device)
    ah, al = a.ptrs
    bh, bl = b.ptrs
    _lib.call("flmm_rotate", a.prec, _lib.ROTATE_UY, ah, al, a.scale_ptr, a.rows_pad, bh, bl, b.scale_ptr,
              b.rows_pad, a.kpad, out.data_ptr(), None, None, None, None, None, _lib.stream())
    return out


def rotate_us(a: Planes, b: Planes) -> tuple[Planes, Planes]:
    """SNP-side scan planes S1 = x, S2 = x^2 of x[i][j] = sum_n a[i][n] b[j][n]."""
    dev = a.hi.device
    s1 = Planes(a.prec, a.rows_pad, b.rows_pad, dev)
    s2 = Planes(a.prec, a.rows_pad, b.rows_pad, dev)
    ah, al = a.ptrs
    bh, bl = b.ptrs
    if a.prec != _lib.PREC_FP32:
        outs = (s1.hi.data_ptr(), s1.lo.data_ptr(), s2.hi.data_ptr(), s2.lo.data_ptr())
    else:
        outs = (s1.hi.data_ptr(), s2.hi.data_ptr(), None, None)
    _lib.call("flmm_rotate", a.prec, _lib.ROTATE_US, ah, al, a.scale_ptr, a.rows_pad, bh, bl, b.scale_ptr,
              b.rows_pad, a.kpad, *outs, s1.scale_ptr, s2.scale_ptr, _lib.stream())
    return s1, s2


def grid_values(n_grid: int, lo: float, hi: float) -> np.ndarray:
    """The log-delta grid exactly as the kernels form it: i * step + lo, last point = hi."""
    if n_grid <= 1:
        return np.array([lo], dtype=np.float64)[:n_grid]
    g = np.arange(n_grid, dtype=np.float64) * ((hi - lo) / (n_grid - 1)) + lo
    g[-1] = hi
    return g


@dataclass
class Factored:
    """Factored weights of one basis and log-delta range (expsum.py): R nodes, padded to rpad, and the SNP-side
    map A (kpad, rpad) FP64, A[k, r] = lam_r exp(-t_r s_k) (zero for k >= neff or r >= R)."""
    R: int
    rpad: int
    t: torch.Tensor    # (R,) FP64
    lam: torch.Tensor  # (R,) FP64
    A: torch.Tensor


def factored_weights(s: torch.Tensor, s_range: tuple[float, float], kpad: int, log_delta_lo: float,
                     log_delta_hi: float) -> Factored | None:
    """Nodes for z = s_k + delta in the basis' eigenvalue range and [exp(lo), exp(hi)]; None when the sum would
    not be shorter than the exact weights (rpad >= kpad)."""
    t, lam = expsum.weight_nodes(*expsum.z_range(s_range[0], s_range[1], log_delta_lo, log_delta_hi))
    R = t.size
    rpad = rup(R, TILE)
    if rpad >= kpad:
        return None
    dev = s.device
    tt = torch.from_numpy(t).to(dev)
    ll = torch.from_numpy(lam).to(dev)
    A = torch.zeros(kpad, rpad, dtype=torch.float64, device=dev)
    A[: s.numel(), :R] = ll.unsqueeze(0) * torch.exp(-s.unsqueeze(1) * tt.unsqueeze(0))
    return Factored(R, rpad, tt, ll, A)


@dataclass
class DeltaResult:
    log_delta: torch.Tensor  # (P,) FP64
    rWr: torch.Tensor        # (P,) FP64
    h2: torch.Tensor         # (P,) FP64
    q1: Planes
    q2: Planes | None        # exact Q2 = w (K = kpad), or with fac the factored Q2 = B (K = fac.rpad)
    idx: torch.Tensor | None = None
    neff: int = 0
    fac: Factored | None = None
    hi_from: int | None = None  # columns from here on have the hi planes only (the screen's); None: all have both
    layout: ScreenLayout | None = None  # the screened columns' bucketed plane layout, or None
    p_pad: int = 0           # rows of the rotated columns' plane (the Q planes have q1.rows_pad >= p_pad)


def attach_factored(prec: int, d: DeltaResult, fac: Factored) -> None:
    """Write the factored Q2 planes B[p][r] = lam_r exp(-t_r delta_p) for d's columns and switch d to them."""
    P = d.log_delta.numel()
    q_rows = d.q1.rows_pad
    q2 = Planes(prec, q_rows, fac.rpad, d.q1.hi.device)
    q2h, q2l = q2.ptrs
    _lib.call("flmm_factored_weight_planes", prec, P, q_rows, fac.rpad, fac.t.data_ptr(), fac.lam.data_ptr(), fac.R,
              d.log_delta.data_ptr(), q2h, q2l, q2.scale_ptr, _hi_from_arg(d.hi_from, q_rows),
              _lib.ptr(d.layout.ws if d.layout else None), _lib.stream())
    d.q2, d.fac = q2, fac


def rotate_snps_f64(Z: torch.Tensor, col_index: torch.Tensor | None, mc: int, mat: torch.Tensor, col0: int,
                    neff: int, rows_pad: int, kpad: int) -> torch.Tensor:
    """US (rows_pad, kpad) FP64, US[i][j] = sum_n Z[n, c(i)] mat[n, col0 + j] (c = col_index, or i), zero outside
    i < mc, j < neff: flmm_rotate_f64 on the FP64 matrix cores."""
    if Z.stride(1) != 1:
        Z = Z.contiguous()
    if mat.stride(1) != 1:
        mat = mat.contiguous()
    if col_index is not None:
        col_index = col_index.to(torch.int64).contiguous()
    out = torch.empty(rows_pad, kpad, dtype=torch.float64, device=Z.device)
    _lib.call("flmm_rotate_f64", Z.data_ptr(), Z.stride(0), _lib.ptr(col_index), mc, Z.shape[0],
              mat.data_ptr() + 8 * col0, mat.stride(0), neff, out.data_ptr(), rows_pad, kpad, _lib.stream())
    return out


def snp_planes_from_f64(prec: int, us: torch.Tensor, mc: int, neff: int,
                        fac: Factored | None) -> tuple[Planes, Planes]:
    """Scan planes of FP64 rotated SNPs us (rows_pad, kpad): S1 = us and S2 = us o us, or S2 = (us o us) A for
    factored weights."""
    rows_pad, kpad = us.shape
    s1 = pack(prec, us, mc, neff, kpad, 1, rows_pad, kpad)
    x2 = us * us
    if fac is None:
        return s1, pack(prec, x2, mc, neff, kpad, 1, rows_pad, kpad)
    c = torch.empty(rows_pad, fac.rpad, dtype=torch.float64, device=us.device)
    _lib.call("flmm_gemm_f64", x2.data_ptr(), rows_pad, kpad, fac.A.data_ptr(), fac.rpad, c.data_ptr(), _lib.stream())
    del x2
    return s1, pack(prec, c, mc, fac.rpad, fac.rpad, 1, rows_pad, fac.rpad)


def factored_snp_planes(a: Planes, bt: Planes, mc: int, fac: Factored) -> tuple[Planes, Planes]:
    """SNP-side planes for factored weights: S1 = x and S2 = (x o x) A, x = the rotated SNPs (rows of a)."""
    s1, x2 = rotate_us_x2(a, bt)
    c = torch.empty(a.rows_pad, fac.rpad, dtype=torch.float64, device=x2.device)
    _lib.call("flmm_gemm_f64", x2.data_ptr(), a.rows_pad, bt.rows_pad, fac.A.data_ptr(), fac.rpad, c.data_ptr(),
              _lib.stream())
    s2 = pack(a.prec, c, mc, fac.rpad, fac.rpad, 1, a.rows_pad, fac.rpad)
    return s1, s2


def delta_grid(uy: torch.Tensor, P: int, p_pad: int, kpad: int, s: torch.Tensor, s_lam: torch.Tensor,
               nco: float, n_grid: int, lo: float, hi: float, col_scale: torch.Tensor | None = None,
               s_min: float = 0.0) -> tuple[torch.Tensor, torch.Tensor]:
    """Per-column grid argmin (int32, (P,)) and the FP64 log-det term at every grid node (the refinement's FP64
    walk to the grid minimum uses it).  With col_scale (power-of-two scales with col_scale[p] |column p| <= 2^14)
    and settings.grid "auto" the grid GEMM runs on split-FP16 MFMA (flmm_delta_grid_split; 0 <= s_min <= every
    s_k), else on FP32 MFMA."""
    if n_grid > TILE:
        raise ValueError(f"n_grid <= {TILE} is supported, got {n_grid}")
    dev = uy.device
    wplane = torch.empty(TILE * kpad, dtype=torch.float32, device=dev)
    lam = torch.empty(TILE, dtype=torch.float32, device=dev)
    lam64 = torch.empty(TILE, dtype=torch.float64, device=dev)
    idx = torch.empty(max(P, 1), dtype=torch.int32, device=dev)
    st = _lib.stream()
    _lib.call("flmm_delta_weights", s.data_ptr(), s.numel(), s_lam.data_ptr(), s_lam.numel(), kpad, n_grid, lo, hi,
              wplane.data_ptr(), lam.data_ptr(), lam64.data_ptr(), st)
    if settings.grid not in ("auto", "fp32"):
        raise ValueError(f"unknown grid setting {settings.grid!r}")
    if col_scale is not None and settings.grid == "auto" and s_min >= 0.0:
        wsplit = torch.empty(4 * TILE * kpad, dtype=torch.uint8, device=dev)  # the grid rows' split planes
        _lib.call("flmm_delta_grid_split", wplane.data_ptr(), lam.data_ptr(), n_grid, float(nco), uy.data_ptr(),
                  col_scale.data_ptr(), float(s_min), lo, hi, P, p_pad, kpad, idx.data_ptr(), wsplit.data_ptr(), st)
    else:
        _lib.call("flmm_delta_grid", wplane.data_ptr(), lam.data_ptr(), n_grid, float(nco), uy.data_ptr(), P, p_pad,
                  kpad, idx.data_ptr(), st)
    return idx[:P], lam64


def delta_grid_idx(uy: torch.Tensor, P: int, p_pad: int, kpad: int, s: torch.Tensor, s_lam: torch.Tensor,
                   nco: float, n_grid: int, lo: float, hi: float) -> torch.Tensor:
    """Per-column FP32 grid argmin (int32, (P,))."""
    return delta_grid(uy, P, p_pad, kpad, s, s_lam, nco, n_grid, lo, hi)[0]


def _hi_from_arg(hi_from: int | None, p_pad: int) -> int:
    return p_pad if hi_from is None else int(hi_from)


def screen_start(prec: int, n_real: int | None, P: int) -> int | None:
    """First column of the certified screen for a scan of P columns with n_real real ones (None: no screen)."""
    if n_real is None or not screen_applies(prec):
        return None
    p0 = rup(n_real, PTILE)
    return p0 if p0 < P else None


SO_NB_MAX = 120   # the bucketed screen's most log-delta buckets
BUCKET_H = 0.2    # its widest bucket (the interpolation's error bound holds up to 1/5)


@dataclass
class ScreenLayout:
    """The bucketed screen's layout of the screened columns: the workspace flmm_delta_refine fills (order), nb
    buckets of width t_h from t_lo, and the Q planes' rows it needs."""
    ws: torch.Tensor
    t_lo: float
    t_h: float
    nb: int
    q_rows: int


def screen_layout(hi_from: int | None, P: int, p_pad: int, t_range: tuple[float, float] | None,
                  dev) -> ScreenLayout | None:
    """The bucketed layout for the screened columns hi_from <= p < P whose log delta lies in t_range (settings.screen
    "auto"), or None where no column is screened, no range is given or it needs more than SO_NB_MAX buckets."""
    if hi_from is None or hi_from >= P or t_range is None or settings.screen != "auto":
        return None
    lo, hi = (float(v) for v in t_range)
    if not (math.isfinite(lo) and math.isfinite(hi)) or hi < lo:
        return None
    nb = max(1, math.ceil((hi - lo) / BUCKET_H))
    while nb <= SO_NB_MAX and (hi - lo) / nb > BUCKET_H:
        nb += 1
    if nb > SO_NB_MAX:
        return None
    t_h = (hi - lo) / nb if hi > lo else BUCKET_H
    lib = _lib.load()
    ws = torch.empty(lib.flmm_screen_order_bytes(P - hi_from), dtype=torch.uint8, device=dev)
    return ScreenLayout(ws, lo, t_h, nb, int(lib.flmm_screen_q_rows(p_pad, hi_from, P)))


def layout_args(layout: ScreenLayout | None, p_pad: int) -> tuple:
    """flmm_delta_refine's (q_rows, t_lo, t_h, nb, order) for a layout (or none)."""
    if layout is None:
        return p_pad, 0.0, 0.0, 0, None
    return layout.q_rows, layout.t_lo, layout.t_h, layout.nb, layout.ws.data_ptr()


def delta_refine(prec: int, mode: int, uy: torch.Tensor, P: int, p_pad: int, kpad: int, s: torch.Tensor,
                 s_lam: torch.Tensor, nco: float, n_grid: int, lo: float, hi: float, idx: torch.Tensor | None,
                 log_delta_in: torch.Tensor | None = None, fac: Factored | None = None,
                 lam64: torch.Tensor | None = None, hi_from: int | None = None,
                 t_range: tuple[float, float] | None = None) -> DeltaResult:
    """Refine (or, for DELTA_FIXED, take log_delta_in) and build the column-side scan planes (with fac, the
    factored Q2 instead of the exact one); lam64: delta_grid's FP64 log-det table for the same grid; hi_from: the
    screen's first column (screen_start), from which only the hi planes are written; t_range (default: (lo, hi)
    for the searching modes, whose fits stay in the grid's range): the range of the columns' log delta, which the
    bucketed screen's layout covers."""
    dev = uy.device
    log_delta = torch.empty(max(P, 1), dtype=torch.float64, device=dev)
    if log_delta_in is not None:
        log_delta[:P].copy_(log_delta_in.to(device=dev, dtype=torch.float64))
    rWr = torch.empty(max(P, 1), dtype=torch.float64, device=dev)
    h2 = torch.empty(max(P, 1), dtype=torch.float64, device=dev)
    if t_range is None and mode != _lib.DELTA_FIXED:
        t_range = (float(lo), float(hi))
    layout = screen_layout(hi_from, P, p_pad, t_range, dev)
    la = layout_args(layout, p_pad)
    q1 = Planes(prec, la[0], kpad, dev)
    q2 = Planes(prec, la[0], kpad, dev) if fac is None else None
    q1h, q1l = q1.ptrs
    q2h, q2l = q2.ptrs if q2 is not None else (None, None)
    _lib.call("flmm_delta_refine", prec, mode, uy.data_ptr(), P, p_pad, kpad, s.data_ptr(), s.numel(),
              s_lam.data_ptr(), s_lam.numel(), float(nco), n_grid, lo, hi, _lib.ptr(idx), _lib.ptr(lam64),
              log_delta.data_ptr(),
              rWr.data_ptr(), h2.data_ptr(), q1h, q1l, q2h, q2l, q1.scale_ptr, q2.scale_ptr if q2 else None,
              _hi_from_arg(hi_from, p_pad), *la, _lib.stream())
    d = DeltaResult(log_delta[:P], rWr[:P], h2[:P], q1, q2, idx, int(s.numel()), hi_from=hi_from, layout=layout,
                    p_pad=p_pad)
    if fac is not None:
        attach_factored(prec, d, fac)
    return d


class _ScanTimer:
    """Optional HIP-event timing of every flmm_wald_scan launch (events on the launch stream)."""

    def __init__(self):
        self.enabled = False
        self.launches: list = []

    def reset(self, enabled: bool = True):
        self.enabled = enabled
        self.launches = []

    def record(self, ev, tests: int, ref_pt: float, alg_pt: float, kernel_pt: float, R: int = 0,
               kind: str = "full") -> None:
        """One launch: its events, Wald tests, and flops per test of the reference algorithm (ref_pt), of the
        launched algorithm (alg_pt) and over the padded K the kernel multiplies (kernel_pt: per product term; the
        split scan issues three such products, the screen one);
        R > 0 for factored weights (the exponential sum's term count); kind: "full" (flmm_wald_scan*), "screen"
        (flmm_wald_screen) or "resolve" (flmm_screen_resolve, recorded with 0 tests)."""
        self.launches.append((ev[0], ev[1], tests, ref_pt, alg_pt, kernel_pt, R, kind))

    def summary(self, kind: str | None = None) -> dict:
        """Launch count, total / mean kernel ms, Wald tests and the three flop totals (ref_flops: the reference
        algorithm, 4 Neff per compat test; alg_flops: what the launches compute, 2 Neff + 2 R per test with
        factored weights; kernel_flops: the same over the padded K), over the launches of one kind or all."""
        torch.cuda.synchronize()
        rows = [r for r in self.launches if kind is None or r[7] == kind]
        ms = [a.elapsed_time(b) for a, b, *_ in rows]
        tot = lambda i: float(sum(r[2] * r[i] for r in rows))  # noqa: E731
        return {"launches": len(ms), "total_ms": float(sum(ms)), "mean_ms": float(np.mean(ms)) if ms else 0.0,
                "tests": int(sum(r[2] for r in rows)), "ref_flops": tot(3), "alg_flops": tot(4),
                "kernel_flops": tot(5), "factored": any(r[6] for r in rows),
                "R": max((r[6] for r in rows), default=0)}


scan_timer = _ScanTimer()


class _StageTimer:
    """Optional HIP-event timing of pipeline stages (``with stage("eigh"): ...``), summed per stage name."""

    def __init__(self):
        self.enabled = False
        self.events: list = []

    def reset(self, enabled: bool = True):
        self.enabled = enabled
        self.events = []

    def summary(self) -> dict:
        torch.cuda.synchronize()
        out: dict = {}
        for name, a, b in self.events:
            out[name] = out.get(name, 0.0) + a.elapsed_time(b)
        return {k: round(v, 3) for k, v in out.items()}


stage_timer = _StageTimer()


@contextlib.contextmanager
def stage(name: str):
    if not stage_timer.enabled:
        yield
        return
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    try:
        yield
    finally:
        b.record()
        stage_timer.events.append((name, a, b))


@dataclass
class ScanOutputs:
    f: torch.Tensor
    beta: torch.Tensor
    se: torch.Tensor
    sfve: torch.Tensor
    nullh2: torch.Tensor
    max_F: torch.Tensor

    @staticmethod
    def allocate(M: int, n_real: int, P: int, dev, *, zero: bool = True) -> "ScanOutputs":
        """zero=False leaves the per-test outputs unset: for callers whose scans write every row."""
        z = lambda: (torch.zeros if zero else torch.empty)(M, n_real, dtype=torch.float64, device=dev)  # noqa: E731
        return ScanOutputs(z(), z(), z(), z(), z(), torch.full((P,), float("-inf"), dtype=torch.float64, device=dev))


def refresh_real_max(out: ScanOutputs, n_real: int) -> None:
    """max_F of the real columns as the column maxima of their (re-scored) per-SNP F."""
    if n_real > 0 and out.f.shape[0] > 0:
        out.max_F[:n_real] = out.f.max(dim=0).values


@dataclass
class ScreenOperands:
    """What the certified screen's FP64 re-evaluation reads: the block's rotated SNPs us (>= mc rows, kpad) FP64,
    the rotated columns' fp32 plane uy, the basis spectrum s (neff) and bounds s_lo <= s <= s_hi."""
    us: torch.Tensor
    uy: torch.Tensor
    s: torch.Tensor
    s_lo: float
    s_hi: float
    neff: int


class _ScreenStats:
    """Optional counts of the screen's FP64 re-evaluations (a device counter, read by summary()) and of its tests,
    those on the sWy-only (bucketed) tile among them."""

    def __init__(self):
        self.counter = None
        self.swy_only = 0
        self.tests = 0
        self.kpad2 = 0

    def reset(self, enabled: bool = True, device=None):
        self.counter = torch.zeros(1, dtype=torch.int64, device=device or "cuda") if enabled else None
        self.swy_only = 0
        self.tests = 0
        self.kpad2 = 0

    def summary(self) -> dict:
        n = int(self.counter.item()) if self.counter is not None else 0
        u = self.swy_only
        return {"screened_tests": self.tests, "fp64_evaluations": n,
                "fraction": n / self.tests if self.tests else 0.0, "swy_only_tests": u,
                "swy_only_fraction": u / self.tests if self.tests else 0.0, "kpad2": self.kpad2}


screen_stats = _ScreenStats()


def screen_applies(prec: int) -> bool:
    """Whether scans of this precision screen their permutation columns (settings.screen)."""
    if settings.screen not in ("auto", "general", "off"):
        raise ValueError(f"unknown screen setting {settings.screen!r}")
    return settings.screen != "off" and prec == _lib.PREC_SPLIT_FP16 and settings.snp_rotation == "fp64"


def wald_scan(s1: Planes, s2: Planes, d: DeltaResult, mc: int, P: int, n_real: int, nu: float,
              row_index: torch.Tensor, out: ScanOutputs, pmax: torch.Tensor | None = None,
              screen: ScreenOperands | None = None) -> None:
    """Scan one SNP block against every column, update out (real-column rows and the running max_F).  With screen
    operands (split-FP16), the column tiles past the real columns run the certified screen instead
    (flmm_wald_scan_screened)."""
    if mc == 0 or P == 0:
        return
    mc_pad = s1.rows_pad
    p_pad = d.q1.rows_pad
    n_mt = (mc + TILE - 1) // TILE
    p_full = P
    if screen is not None and s1.prec == _lib.PREC_SPLIT_FP16 and mc <= 64 * TILE:
        p_full = min(P, rup(n_real, PTILE))
    if d.hi_from is not None and d.hi_from < p_full:
        raise ValueError(f"wald_scan: the column planes past column {d.hi_from} hold the hi halves only (the screen's), "
                         f"but columns up to {p_full} need the full split")
    need = n_mt * p_pad
    if pmax is None or pmax.numel() < need:
        pmax = torch.empty(need, dtype=torch.float64, device=s1.hi.device)
    if p_full < P:
        _screened(s1, s2, d, mc, p_full, P, nu, out, pmax, screen)
    if p_full == 0:
        return
    P = p_full
    s1h, s1l = s1.ptrs
    s2h, s2l = s2.ptrs
    q1h, q1l = d.q1.ptrs
    q2h, q2l = d.q2.ptrs
    real = n_real > 0
    o = (out.f, out.beta, out.se, out.sfve, out.nullh2)
    optr = [t.data_ptr() if real else None for t in o]
    st = _lib.stream()
    ev = None
    if scan_timer.enabled:
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        ev[0].record()
    if d.fac is not None:  # factored weights: sWs over the first fac.rpad entries (S2 / Q2 hold K = rpad)
        if s2.kpad != d.fac.rpad or d.q2.kpad != d.fac.rpad:
            raise ValueError("factored scan: the SNP planes were not built for these weights")
        _lib.call("flmm_wald_scan_lowrank", s1.prec, s1h, s1l, s2h, s2l, s1.scale_ptr, s2.scale_ptr, mc_pad, q1h, q1l,
                  q2h, q2l, d.q1.scale_ptr, d.q2.scale_ptr, p_pad, s1.kpad, d.fac.rpad, mc, P, n_real,
                  d.rWr.data_ptr(), d.h2.data_ptr(), float(nu), row_index.data_ptr(), *optr, pmax.data_ptr(), st)
    else:
        _lib.call("flmm_wald_scan", s1.prec, s1h, s1l, s2h, s2l, s1.scale_ptr, s2.scale_ptr, mc_pad, q1h, q1l, q2h,
                  q2l, d.q1.scale_ptr, d.q2.scale_ptr, p_pad, s1.kpad, mc, P, n_real, d.rWr.data_ptr(),
                  d.h2.data_ptr(), float(nu), row_index.data_ptr(), *optr, pmax.data_ptr(), st)
    if ev is not None:
        ev[1].record()
        k2, r2 = (d.fac.rpad, d.fac.R) if d.fac else (s1.kpad, d.neff)
        scan_timer.record(ev, mc * P, 4.0 * d.neff, 2.0 * (d.neff + r2), 2.0 * (s1.kpad + k2), d.fac.R if d.fac else 0)
    _lib.call("flmm_max_reduce", pmax.data_ptr(), n_mt, p_pad, P, out.max_F.data_ptr(), st)


def _screened(s1: Planes, s2: Planes, d: DeltaResult, mc: int, p_first: int, P: int, nu: float, out: ScanOutputs,
              rec: torch.Tensor, screen: ScreenOperands) -> None:
    """flmm_wald_scan_screened over the columns p_first <= p < P of one SNP block (rec: >= ceil(mc/128) x
    (P - p_first) 8-byte words of workspace)."""
    kpad2 = d.fac.rpad if d.fac is not None else s1.kpad
    us = screen.us
    cnt = screen_stats.counter
    # the bucketed screen: the layout the refinement wrote for exactly these columns
    order = d.layout.ws if (d.layout is not None and d.hi_from == p_first) else None
    work = None
    if order is not None:
        work = torch.empty(_lib.load().flmm_screen_work_bytes(s1.rows_pad, s1.kpad), dtype=torch.uint8,
                           device=rec.device)
    args = (s1.hi.data_ptr(), s2.hi.data_ptr(), s1.scale_ptr, s2.scale_ptr, s1.rows_pad, d.q1.hi.data_ptr(),
            d.q2.hi.data_ptr(), d.q1.scale_ptr, d.q2.scale_ptr, d.p_pad or d.q1.rows_pad, d.q1.rows_pad, s1.kpad,
            kpad2, mc, p_first, P, us.data_ptr(), us.stride(0), screen.uy.data_ptr(), screen.s.data_ptr(),
            screen.neff, float(screen.s_lo), float(screen.s_hi), d.log_delta.data_ptr(), d.rWr.data_ptr(), float(nu),
            out.max_F.data_ptr(), rec.data_ptr(), _lib.ptr(order), _lib.ptr(work), _lib.ptr(cnt), _lib.stream())
    if cnt is not None:
        screen_stats.tests += mc * (P - p_first)
        screen_stats.kpad2 = kpad2
        if order is not None:  # every test of the bucketed layout runs on the sWy-only tile
            screen_stats.swy_only += mc * (P - p_first)
    if not scan_timer.enabled:
        _lib.call("flmm_wald_scan_screened", *args)
        return
    # timed: the screen and the resolution as two calls, each between HIP events on the launch stream
    k2, r2 = (d.fac.rpad, d.fac.R) if d.fac else (s1.kpad, d.neff)
    for name, kind in (("flmm_wald_screen", "screen"), ("flmm_screen_resolve", "resolve")):
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        ev[0].record()
        _lib.call(name, *args)
        ev[1].record()
        # the resolution adds time, not tests: the tests count once, with the screen
        scan_timer.record(ev, mc * (P - p_first) if kind == "screen" else 0, 4.0 * d.neff, 2.0 * (d.neff + r2),
                          2.0 * (s1.kpad + k2), d.fac.R if d.fac else 0, kind=kind)
