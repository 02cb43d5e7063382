"""FP64 CPU oracle for the low-rank scan path -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the reference's exact low-rank spectral LMM (fasterlmm/lowrank.py), its streamed LOCO
loop (fasterlmm/extreme_scan.py:28-77) and the randomised range finder (fasterlmm/extreme.py:23-45).  Only
tests/ may import it, as the checker of the HIP path.

Pinning (checked independently of this restatement in tests/test_oracle_lowrank.py):
  lowrank_basis            U_eff orthonormal and orthogonal to X; U diag(s) U^T equals the projected kinship
                           (I - X X^+) G G^T (I - X X^+) / M_kin; the k >= N branch gives the same spectrum
  rotate_phenos            U UY + UUY = Y - X X^+ Y, U^T UUY = 0
  profile loss / delta     the dense -2 log-likelihood of N(0, K_r + delta I) restricted to null(X^T), i.e. the
                           compat loss on the full projected spectrum with the null eigenvalues at 0
  snp_wald_scan_lowrank    equals the full-rank compat scan (oracle.lmm, pinned by the reference's golden vectors)
                           on K = G G^T / M_kin at a common log delta: the reference's own cross-check
                           (tests/test_lowrank.py:124-151)
  loco / single_k          per-chromosome restatement (tests/test_lowrank.py:154-187)
  randomized_eigh          exact on matrices of rank <= rank + n_oversample; top eigenpairs of a spectrum with
                           a gap match numpy eigh
"""
from __future__ import annotations

import numpy as np

from oracle.lmm import LOG2PI, _empty_result, _golden_refine, compat_wald_from_sums, pinv, pos_floor


def lowrank_basis(G, X, *, sv_floor=1e-10):
    """lowrank_basis (lowrank.py:59-96): regress X out of the kinship factor G (N, k_in), eigh of the k_in x k_in
    Gram when k_in < N (U = G_ V / sv, s = sv^2 / k_in, singular values <= sv_floor dropped), else eigh of
    G_ G_^T / k_in with eigenvalues <= sv_floor^2 / k_in dropped.  Returns dict(s, U_eff, X, Xpinv, Neff)."""
    N, k_in = G.shape
    D = X.shape[1]
    Xpinv = pinv(X)
    G_ = G - X @ (Xpinv @ G)
    if k_in < N:
        evals, V = np.linalg.eigh(G_.T @ G_)
        sv = np.sqrt(np.clip(evals, 0.0, None))
        keep = sv > sv_floor
        sv, V = sv[keep], V[:, keep]
        U_eff = G_ @ (V / sv)
        s = sv * sv / k_in
    else:
        evals, U = np.linalg.eigh((G_ @ G_.T) / k_in)
        keep = np.clip(evals, 0.0, None) > sv_floor * sv_floor / k_in
        s = np.clip(evals[keep], 0.0, None)
        U_eff = np.ascontiguousarray(U[:, keep])
    return {"s": s, "U_eff": U_eff, "X": X, "Xpinv": Xpinv, "Neff": N - D}


def rotate_phenos(basis, Y):
    """rotate_phenos (lowrank.py:99-111): Y_r = Y - X X^+ Y, UY = U^T Y_r, UUY = Y_r - U UY."""
    Y_r = Y - basis["X"] @ (basis["Xpinv"] @ Y)
    UY = basis["U_eff"].T @ Y_r
    return UY, Y_r - basis["U_eff"] @ UY


def lowrank_rotate(G, X, Y, *, sv_floor=1e-10):
    """lowrank_rotate (lowrank.py:114-122)."""
    b = lowrank_basis(G, X, sv_floor=sv_floor)
    UY, UUY = rotate_phenos(b, Y)
    return {**b, "UY": UY, "UUY": UUY}


def profile_loss_lowrank(log_delta, s, UY, yy_off, Neff):
    """(G, P) low-rank -2 log-likelihood (lowrank.py:128-150): on-rank weights 1/(s + delta), off-rank residual
    sum of squares over delta, log det sum log(s + delta) + (Neff - k) log delta."""
    log_delta = np.atleast_1d(np.asarray(log_delta, dtype=np.float64))
    k = s.shape[0]
    delta = np.exp(log_delta)
    Sd = s[None, :] + delta[:, None]
    rWr = np.maximum((1.0 / Sd) @ (UY * UY) + yy_off[None, :] / delta[:, None], pos_floor())
    logdet = np.log(Sd).sum(axis=1) + (Neff - k) * np.log(delta)
    return Neff * (LOG2PI + np.log(rWr / Neff) + 1.0) + logdet[:, None]


def profile_loss_lowrank_per_pheno(log_delta, s, UY, yy_off, Neff):
    """(P,) low-rank loss, one log delta per column (lowrank.py:153-169)."""
    k = s.shape[0]
    delta = np.exp(np.asarray(log_delta, dtype=np.float64))
    Sd = s[None, :] + delta[:, None]
    rWr = np.maximum((UY.T * UY.T / Sd).sum(axis=1) + yy_off / delta, pos_floor())
    logdet = np.log(Sd).sum(axis=1) + (Neff - k) * np.log(delta)
    return Neff * (LOG2PI + np.log(rWr / Neff) + 1.0) + logdet


def fit_delta_grid_lowrank(spec, *, n_grid=64, log_delta_min=-10.0, log_delta_max=10.0, refine=True):
    """fit_delta_grid_lowrank (lowrank.py:172-208): grid then 50 golden-section steps."""
    s, UY, Neff = spec["s"], spec["UY"], spec["Neff"]
    yy_off = (spec["UUY"] * spec["UUY"]).sum(axis=0)
    grid = np.linspace(log_delta_min, log_delta_max, n_grid)
    loss = profile_loss_lowrank(grid, s, UY, yy_off, Neff)
    if not refine:
        return grid[np.argmin(loss, axis=0)]
    return _golden_refine(grid, loss, lambda ld: profile_loss_lowrank_per_pheno(ld, s, UY, yy_off, Neff),
                          log_delta_min, log_delta_max)


def snp_wald_scan_lowrank(spec, log_delta, S, *, n_real=None):
    """snp_wald_scan_lowrank (lowrank.py:239-316): S_r = S - X X^+ S, US = U^T S_r, UUS = S_r - U US;
    sWs = w.US^2 + |UUS|^2 / delta, sWy = (w o UY)^T US + UUS^T UUY / delta, F with dof Neff - 1."""
    s, UY, UUY, X, Xpinv, U = spec["s"], spec["UY"], spec["UUY"], spec["X"], spec["Xpinv"], spec["U_eff"]
    Neff = spec["Neff"]
    P, M = UY.shape[1], S.shape[1]
    n_real = P if n_real is None else min(n_real, P)
    delta = np.exp(log_delta)
    w = 1.0 / (s[None, :] + delta[:, None])  # (P, k)
    yy_off = (UUY * UUY).sum(axis=0)
    rWr = np.maximum((w * UY.T * UY.T).sum(axis=1) + yy_off / delta, pos_floor())
    S_r = S - X @ (Xpinv @ S)
    US = U.T @ S_r
    UUS = S_r - U @ US
    ss_off = (UUS * UUS).sum(axis=0)
    sWs = w @ (US * US) + ss_off[None, :] / delta[:, None]
    sWy = (w * UY.T) @ US + (UUS.T @ UUY).T / delta[:, None]
    f, beta, se, sfve = compat_wald_from_sums(sWy, sWs, rWr[:, None], Neff - 1)
    out = _empty_result(M, n_real, P)
    out["max_F"] = f.max(axis=1) if M > 0 else np.full(P, -np.inf)
    out["f"] = f[:n_real].T.copy()
    out["beta"] = beta[:n_real].T.copy()
    out["se"] = se[:n_real].T.copy()
    out["sfve"] = sfve[:n_real].T.copy()
    h2 = 1.0 / (1.0 + delta[:n_real])
    out["nullh2"] = np.broadcast_to(h2[None, :], (M, n_real)).copy()
    return out


def single_k_scan_lowrank(G, Z, X, Y, *, n_real=None, n_grid=64):
    """single_k_scan_lowrank (lowrank.py:363-373)."""
    spec = lowrank_rotate(G, X, Y)
    return snp_wald_scan_lowrank(spec, fit_delta_grid_lowrank(spec, n_grid=n_grid), Z, n_real=n_real)


def loco_scan_lowrank(G, g_chrom, Z, z_chrom, X, Y, *, n_real=None, n_grid=64, return_log_delta=False):
    """loco_scan_lowrank (lowrank.py:322-360) and the streamed/resident LOCO core (extreme_scan.py:28-77):
    per chromosome c the kinship factor drops c's markers, the basis, delta fit and scan of c's SNPs."""
    M, P = Z.shape[1], Y.shape[1]
    nr = P if n_real is None else min(n_real, P)
    out = _empty_result(M, nr, P)
    lds = {}
    for c in sorted(np.unique(z_chrom).tolist()):
        spec = lowrank_rotate(G[:, g_chrom != c], X, Y)
        ld = fit_delta_grid_lowrank(spec, n_grid=n_grid)
        lds[c] = ld
        test = z_chrom == c
        r = snp_wald_scan_lowrank(spec, ld, Z[:, test], n_real=nr)
        for k in ("f", "beta", "se", "sfve", "nullh2"):
            out[k][test] = r[k]
        out["max_F"] = np.maximum(out["max_F"], r["max_F"])
    return (out, lds) if return_log_delta else out


def randomized_eigh(Z, rank, Omega, *, n_iter=2):
    """randomized_eigh (extreme.py:23-45) with the Gaussian probe Omega (N, rank + oversample) given: Y = K Omega
    with K = Z Z^T / M, n_iter subspace iterations, Q = qr(Y), eigh of Q^T K Q, top `rank` pairs descending."""
    M = Z.shape[1]
    Y = Z @ (Z.T @ Omega) / M
    for _ in range(n_iter):
        Y = Z @ (Z.T @ Y) / M
    Q, _ = np.linalg.qr(Y)
    B = Q.T @ (Z @ (Z.T @ Q)) / M
    s_small, U_small = np.linalg.eigh(B)
    s_small = np.clip(s_small, 0.0, None)
    idx = np.argsort(-s_small, kind="stable")[:rank]
    return s_small[idx], Q @ U_small[:, idx]
