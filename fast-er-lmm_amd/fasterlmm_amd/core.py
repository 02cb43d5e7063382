"""Drop-in spectral LMM core on MI355X: the reference's fasterlmm/core.py API (names, signatures, dataclasses,
error behaviour) with the scan, the rotations and the delta search running in the gfx950 kernels.

Reference: fasterlmm/core.py (gbrach/FaST-ER-LMM v1.6.2).

* Inputs may live on the CPU or the GPU; the work always runs on the GPU and the outputs come back on the
  device of the input, in FP64, as in the reference (core.py:322-326, 701-706).  There is no CPU fallback:
  without a ROCm GPU or without the built library every entry point raises.
* ``snp_chunk`` / ``pheno_chunk`` are accepted for signature compatibility; tiling is internal (128 x 128 MFMA
  tiles) and, as in the reference (chunking-invariance tests), invisible in the results.
* Precision: the rotations and the scan products run as split-BF16 MFMA (hi + lo, three products, FP32
  accumulation) by default, or FP32 MFMA (``engine.configure(precision="fp32")``); every reduction into F,
  beta, SE, the delta search and the projections are FP64.
* Delta refinement defaults to safeguarded Newton to the local minimiser inside the reference's golden-section
  bracket; ``engine.configure(delta="golden")`` replays the reference's 50 golden-section steps instead.
* The OLS projection of the test SNPs, S - X X^+ S (core.py:628), is not formed: every kept eigenvector is
  orthogonal to the column space of X, so U_eff^T (S - X X^+ S) = U_eff^T S exactly in exact arithmetic.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch
from torch import Tensor

from . import _lib, engine
from .engine import TILE, KCHUNK, Planes, rup


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
    return ScanResult(*(getattr(res, k).to(home) for k in ("f", "beta", "se", "sfve", "nullh2", "max_F")))


def _eigh(K: Tensor) -> tuple[Tensor, Tensor]:
    """FP64 symmetric eigendecomposition on the GPU (rocSOLVER through torch), ascending (core.py:54-77).  An
    out-of-memory error retries on the host LAPACK like the reference's fallback; any other error propagates."""
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


def eigendecompose(K: Tensor, jitter: float = 1e-6) -> tuple[Tensor, Tensor]:
    """eigh(K + jitter I) with eigenvalues clamped at 0 (core.py:89-102)."""
    home, (Kd,) = _dev(K)
    N = Kd.shape[0]
    s, U = _eigh(Kd + jitter * torch.eye(N, dtype=Kd.dtype, device=Kd.device))
    s = torch.clamp(s, min=0.0)
    return s.to(home), U.to(home)


# COMPAT PATH -------

class _CompatGPU:
    """Device-side state of one projected eigendecomposition: U_eff^T planes and the rotated columns."""

    def __init__(self, s: Tensor, U: Tensor, D: int, prec: int):
        N = U.shape[0]
        self.N, self.D, self.neff = N, D, N - D
        self.npad = rup(N, KCHUNK)
        self.kpad = rup(max(self.neff, 1), TILE)
        self.prec = prec
        self.s = s.contiguous()
        # U_eff^T planes: row j = eigenvector D + j, element k = U[k, D + j]
        self.ueff = engine.pack(prec, U, self.neff, N, 1, N, self.kpad, self.npad, offset=D)
        self.uy = None  # fp32 plane of the rotated columns
        self.P = 0
        self.p_pad = 0
        self.delta = None  # last DeltaResult (its log_delta tensor identifies it)

    def rotate_columns(self, Y_r: Tensor) -> None:
        """UY plane of U_eff^T Y_r for Y_r (N, P)."""
        N, P = Y_r.shape
        self.P, self.p_pad = P, rup(max(P, 1), TILE)
        ypl = engine.pack(self.prec, Y_r.contiguous(), P, N, 1, P, self.p_pad, self.npad)
        self.uy = engine.rotate_uy(ypl, self.ueff)

    def uy_dense(self) -> Tensor:
        return engine.fp32_plane_to_dense(self.uy, self.p_pad, self.kpad, self.P, self.neff).T.double()

    def set_uy_dense(self, UY: Tensor) -> None:
        """Rebuild the UY plane from a caller-supplied (Neff, P) FP64 matrix."""
        neff, P = UY.shape
        self.P, self.p_pad = P, rup(max(P, 1), TILE)
        pl = engine.pack(_lib.PREC_FP32, UY.contiguous(), P, neff, 1, P, self.p_pad, self.kpad)
        self.uy = pl.hi

    def fit(self, n_grid: int, lo: float, hi: float, refine: bool, mode: int | None = None):
        idx = engine.delta_grid_idx(self.uy, self.P, self.p_pad, self.kpad, self.s, self.s, float(self.neff),
                                    n_grid, lo, hi)
        if not refine:
            grid = torch.from_numpy(engine.grid_values(n_grid, lo, hi)).to(self.s.device)
            return grid[idx.long()], None
        d = engine.delta_refine(self.prec, engine.delta_code() if mode is None else mode, self.uy, self.P,
                                self.p_pad, self.kpad, self.s, self.s, float(self.neff), n_grid, lo, hi, idx)
        self.delta = d
        return d.log_delta, d

    def delta_for(self, log_delta: Tensor) -> engine.DeltaResult:
        d = self.delta
        if d is not None and d.log_delta.data_ptr() == log_delta.data_ptr() and d.q1.prec == self.prec:
            return d
        return engine.delta_refine(self.prec, _lib.DELTA_FIXED, self.uy, self.P, self.p_pad, self.kpad, self.s,
                                   self.s, float(self.neff), 1, 0.0, 0.0, None, log_delta_in=log_delta)

    def scan(self, Z: Tensor, row_index: Tensor, n_real: int, out: engine.ScanOutputs, d: engine.DeltaResult,
             block: int = 8192) -> None:
        """Scan the SNP columns Z[:, row_index] (Z is (N, M_total) FP64) into out."""
        M_tot = Z.shape[1]
        Mc = row_index.numel()
        for b0 in range(0, Mc, block):
            ri = row_index[b0:b0 + block]
            mc = ri.numel()
            mc_pad = rup(mc, TILE)
            spl = engine.pack(self.prec, Z, mc, self.N, 1, M_tot, mc_pad, self.npad, row_index=ri)
            s1, s2 = engine.rotate_us(spl, self.ueff)
            del spl
            engine.wald_scan(s1, s2, d, mc, self.P, n_real, float(self.neff - 1), ri, out)


def _compat_project(K: Tensor, X: Tensor) -> tuple[Tensor, Tensor, Tensor]:
    """Xpinv, and the projected eigendecomposition of (I - X X^+)(K + I)(I - X X^+) (core.py:495-510)."""
    N = K.shape[0]
    Xpinv = torch.linalg.pinv(X)
    Iperp = torch.eye(N, dtype=K.dtype, device=K.device) - X @ Xpinv
    Kp = (Iperp @ (K + torch.eye(N, dtype=K.dtype, device=K.device))) @ Iperp
    s_full, U = _eigh(Kp)
    return Xpinv, s_full, U


def _spec_state(spec: CompatSpectrum) -> _CompatGPU:
    st = getattr(spec, "_gpu", None)
    if st is None:
        raise ValueError("CompatSpectrum was not built by fasterlmm_amd.core.fastlmm_compat_rotate")
    return st


def fastlmm_compat_rotate(K: Tensor, X: Tensor, Y: Tensor) -> CompatSpectrum:
    """Projected rotation (core.py:487-519): eigh of (I - X X^+)(K + I)(I - X X^+), drop the first D = ncols(X)
    eigenpairs, s = clamp(lambda - 1, 0), UY = U_eff^T (Y - X X^+ Y) (the product on MFMA)."""
    home, (Kd, Xd, Yd) = _dev(K, X, Y)
    N = Kd.shape[0]
    D = Xd.shape[1]
    Xpinv, s_full, U = _compat_project(Kd, Xd)
    s = torch.clamp(s_full[D:N] - 1.0, min=0.0)
    st = _CompatGPU(s, U, D, engine.prec_code())
    Y_r = Yd - Xd @ (Xpinv @ Yd)
    st.rotate_columns(Y_r)
    spec = CompatSpectrum(s=s.to(home), U_eff=U[:, D:N].contiguous().to(home), X=Xd.to(home), Xpinv=Xpinv.to(home),
                          UY=st.uy_dense().to(home))
    spec._gpu = st
    return spec


def fit_delta_grid_compat(spec: CompatSpectrum, *, n_grid: int = 64, log_delta_min: float = -10.0,
                          log_delta_max: float = 10.0, refine: bool = True) -> Tensor:
    """Per-column log delta for the compat path: grid argmin then refinement in [g[i-1], g[i+1]]
    (core.py:561-595).  Returns (P,) FP64 on the spectrum's device."""
    st = _spec_state(spec)
    home = spec.s.device
    ld, _ = st.fit(n_grid, float(log_delta_min), float(log_delta_max), refine)
    return ld.to(home)


def snp_wald_scan_compat(spec: CompatSpectrum, log_delta: Tensor, S: Tensor, *, snp_chunk: int = 4096,
                         pheno_chunk: int = 256, n_real: int | None = None) -> ScanResult:
    """Wald F / beta / SE / SnpFractVarExpl of every SNP column of S against every column of the spectrum, in the
    projected basis with dof Neff - 1 (core.py:601-671).  Real columns (the first n_real) keep full detail, every
    column keeps its maximum F."""
    st = _spec_state(spec)
    home, (Sd, ldd) = _dev(S, log_delta)
    P = st.P
    M = Sd.shape[1]
    n_real = P if n_real is None else min(int(n_real), P)
    d = st.delta_for(ldd)
    out = engine.ScanOutputs.allocate(M, n_real, P, Sd.device)
    st.scan(Sd.contiguous(), torch.arange(M, device=Sd.device, dtype=torch.int64), n_real, out, d)
    return _home(ScanResult(out.f, out.beta, out.se, out.sfve, out.nullh2, out.max_F), home)


def _chrom_labels(chrom) -> np.ndarray:
    return np.asarray(chrom)


def loco_scan_compat(Z: Tensor, X: Tensor, Y: Tensor, chrom: np.ndarray, *, n_real: int | None = None,
                     on_chrom=None, Z_kin: Tensor | None = None, chrom_kin: np.ndarray | None = None,
                     n_grid: int = 64) -> ScanResult:
    """Leave-one-chromosome-out compat scan (core.py:677-726).  For each chromosome c: K from every other
    chromosome's SNPs (of Z, or of the kinship panel Z_kin), the projected eigendecomposition, per-column delta,
    and the scan of c's SNPs; max_F is the genome-wide max over chromosomes.  Z must be pre-standardised."""
    if (Z_kin is None) != (chrom_kin is None):
        raise ValueError("Z_kin and chrom_kin go together")
    home, (Zd, Xd, Yd, Zkd) = _dev(Z, X, Y, Z_kin)
    chrom = _chrom_labels(chrom)
    N, M = Zd.shape
    P = Yd.shape[1]
    D = Xd.shape[1]
    n_real = P if n_real is None else min(int(n_real), P)
    dev = Zd.device
    prec = engine.prec_code()
    out = engine.ScanOutputs.allocate(M, n_real, P, dev)
    Zk = Zd if Zkd is None else Zkd
    ck = chrom if chrom_kin is None else _chrom_labels(chrom_kin)
    K_all = Zk @ Zk.T
    Xpinv = torch.linalg.pinv(Xd)
    eye = torch.eye(N, dtype=torch.float64, device=dev)
    Iperp = eye - Xd @ Xpinv
    Y_r = (Yd - Xd @ (Xpinv @ Yd)).contiguous()
    npad = rup(N, KCHUNK)
    p_pad = rup(max(P, 1), TILE)
    ypl = engine.pack(prec, Y_r, P, N, 1, P, p_pad, npad)
    del Y_r
    chroms = sorted(np.unique(chrom).tolist())
    for k, c in enumerate(chroms):
        test = torch.from_numpy(np.flatnonzero(chrom == c)).to(dev)
        kin_c = torch.from_numpy(np.flatnonzero(ck == c)).to(dev)
        Zc = Zk.index_select(1, kin_c)
        K = (K_all - Zc @ Zc.T) / float(Zk.shape[1] - kin_c.numel())
        del Zc
        Kp = (Iperp @ (K + eye)) @ Iperp
        s_full, U = _eigh(Kp)
        s = torch.clamp(s_full[D:N] - 1.0, min=0.0)
        st = _CompatGPU(s, U, D, prec)
        st.P, st.p_pad = P, p_pad
        st.uy = engine.rotate_uy(ypl, st.ueff)
        _, d = st.fit(n_grid, -10.0, 10.0, True)
        st.scan(Zd, test, n_real, out, d)
        if on_chrom is not None:
            on_chrom(k + 1, len(chroms))
    return _home(ScanResult(out.f, out.beta, out.se, out.sfve, out.nullh2, out.max_F), home)


def single_k_scan_compat(Z: Tensor, X: Tensor, Y: Tensor, *, n_real: int | None = None,
                         Z_kin: Tensor | None = None, n_grid: int = 64) -> ScanResult:
    """One kinship over all SNPs (or the Z_kin panel), compat rotation, delta fit, scan of every SNP
    (core.py:729-740)."""
    from .io import grm
    home, (Zd, Zkd) = _dev(Z, Z_kin)
    K = grm(Zd if Zkd is None else Zkd)
    spec = fastlmm_compat_rotate(K, X, Y)
    ld = fit_delta_grid_compat(spec, n_grid=n_grid)
    return _home(snp_wald_scan_compat(spec, ld, Zd, n_real=n_real), home)
