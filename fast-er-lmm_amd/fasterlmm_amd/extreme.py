"""Randomised top-r eigendecomposition (the reference's fasterlmm/extreme.py:23-45) on the GPU.

Halko, Martinsson & Tropp (SIAM Review 2011), algorithm 4.4 (randomised subspace iteration) + 5.3 (direct
eigenvalue decomposition): Y = K Omega for a Gaussian probe Omega (N, r + oversample), n_iter subspace iterations
Y <- K Y, Q = qr(Y), eigh of the small Q^T K Q, U = Q U_small; the top r pairs are kept.  K is only ever applied to
a block of vectors -- K = Z Z^T / M as Z (Z^T Y) / M -- so the N x N kernel is never formed, and the tall
orthonormalisations are Gram-based (orthonormal_basis): GEMMs and small eighs, no Householder QR of the N x b
block (rocSOLVER's takes ~1.5 s at 50,000 x 2,020, the GEMMs ~0.1 s).  FP64 unless stated.
"""
from __future__ import annotations

import torch
from torch import Tensor


def orthonormal_basis(Y: Tensor, *, passes: int = 2, rtol: float = 1e-13) -> Tensor:
    """An orthonormal basis (FP64) of span(Y) for a tall Y (N, b), by CholeskyQR2: Q = Y L^-T with L the Cholesky
    factor of the small Gram Y^T Y (inverted explicitly, applied as a GEMM), twice (the first pass leaves an orthogonality error ~ eps cond(Y)^2, the second
    removes it) -- GEMMs, one b x b Cholesky and one triangular solve per pass instead of a Householder QR of the
    tall matrix.  Where a Gram is numerically singular (Y rank-deficient) that pass takes its eigendecomposition
    instead and drops the directions with eigenvalue <= rtol x the largest."""
    Q = Y.double()
    for _ in range(passes):
        G = Q.T @ Q
        G = 0.5 * (G + G.T)
        L, info = torch.linalg.cholesky_ex(G)
        lam_min = None
        if int(info.item()) == 0:
            d = torch.diagonal(L)
            lam_min = float((d.min() / d.max()).item()) ** 2  # a cheap lower estimate of the Gram's conditioning
        if lam_min is not None and lam_min > 1e3 * rtol:
            # invert the small factor and apply it with one GEMM (a triangular solve against the tall block would
            # need a workspace of its size)
            Li = torch.linalg.solve_triangular(L, torch.eye(L.shape[0], dtype=L.dtype, device=L.device), upper=False)
            Q = Q @ Li.T
        else:
            lam, V = torch.linalg.eigh(G)
            keep = lam > rtol * lam[-1].clamp(min=0.0)
            Q = Q @ (V[:, keep] / lam[keep].sqrt())
    return Q


def randomized_range(apply_k, N: int, rank: int, *, n_oversample: int = 20, n_iter: int = 2, seed: int = 19930909,
                     dtype=torch.float64, device=None) -> tuple[Tensor, Tensor]:
    """Top-`rank` eigenpairs (descending) of the symmetric PSD operator apply_k (an (N, b) -> (N, b) product)."""
    target = min(rank + n_oversample, N)
    gen = torch.Generator(device=device).manual_seed(seed)
    Omega = torch.randn(N, target, generator=gen, device=device, dtype=dtype)
    Y = apply_k(Omega)
    for _ in range(n_iter):
        # re-orthonormalise between iterations (keeps the subspace well conditioned; same span)
        Y = apply_k(orthonormal_basis(Y).to(dtype))
    Q = orthonormal_basis(Y).to(dtype)
    B = Q.T @ apply_k(Q)
    B = 0.5 * (B + B.T)
    s_small, U_small = torch.linalg.eigh(B)
    s_small = torch.clamp(s_small, min=0.0)
    idx = torch.argsort(s_small, descending=True)[:rank]
    return s_small[idx], Q @ U_small[:, idx]


def randomized_eigh(Z: Tensor, rank: int = 200, *, n_oversample: int = 20, n_iter: int = 2,
                    seed: int = 19930909) -> tuple[Tensor, Tensor]:
    """Top-r eigendecomposition of K = Z Z^T / M (extreme.py:23-45): returns (s_top (r,) descending, U_top (N, r)),
    on the device and dtype of Z's GPU copy (FP64)."""
    from .core import _dev
    home, (Zd,) = _dev(Z)
    N, M = Zd.shape
    s, U = randomized_range(lambda Y: Zd @ (Zd.T @ Y) / M, N, int(rank), n_oversample=n_oversample,
                            n_iter=n_iter, seed=seed, dtype=Zd.dtype, device=Zd.device)
    return s.to(home), U.to(home)
