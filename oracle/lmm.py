"""FP64 CPU oracle for the LMM association scan -- TEST INFRASTRUCTURE ONLY.

This module is a numpy restatement of the reference's spectral LMM (fasterlmm/core.py, fasterlmm/io.py).
It exists to check the HIP path.  Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or the timed CPU baseline -- never as the product path.

Every function cites the reference file:line it restates.  The arithmetic is kept in the same order as the
reference where that order decides rounding-visible behaviour (clamps, floors, the golden-section update),
so that results agree to ~1e-12 relative, not just in distribution.

Pinning (what checks each function independently of this restatement; see tests/test_oracle_*.py):
  standardise_columns   closed form + the reference's own io tests (NaN, constant column, ddof 0)
  grm                   brute force Z Z^T / M
  eigendecompose        numpy eigh invariants (ascending, orthonormal, reconstruction)
  compat_rotate         drop-D shapes, U_eff orthonormal, UY == U_eff^T (Y - X X^+ Y), golden vectors
  profile losses        dense slogdet(K + delta I) / GLS closed forms
  fit_delta_grid(*)     grid argmin exactness (refine=False lands on a node), bounded scalar minimiser,
                        golden vectors from the reference implementation
  snp_wald_scan(*)      textbook dense GLS Wald test per (SNP, column), OLS limit delta -> inf,
                        golden vectors from the reference implementation
  loco/single_k         per-chromosome restatement, golden vectors
"""
from __future__ import annotations

import numpy as np

LOG2PI = float(np.log(2.0 * np.pi))
INV_PHI = 2.0 / (1.0 + 5.0 ** 0.5)
INV_PHI2 = INV_PHI * INV_PHI


def pos_floor(dtype=np.float64) -> float:
    """Clamp floor per dtype (core.py:80-86): 1e-300 for float64, 1e-30 for float32."""
    return 1e-300 if np.dtype(dtype) == np.float64 else 1e-30


# --------------------------------------------------------------------------------------------------------
# io.py restatements
# --------------------------------------------------------------------------------------------------------

def standardise_columns(arr: np.ndarray) -> np.ndarray:
    """Unit standardiser (io.py:247-254): NaN-aware mean and population std per column, (x-mu)/sd,
    sd == 0 -> inf (the column becomes 0), NaN -> 0 after standardising."""
    a = np.asarray(arr, dtype=np.float64).copy()
    nan_mask = np.isnan(a)
    col_mean = np.nanmean(a, axis=0)
    col_std = np.nanstd(a, axis=0)
    col_std = np.where(col_std == 0, np.inf, col_std)
    a = (a - col_mean) / col_std
    a[nan_mask] = 0.0
    return a


def grm(Z: np.ndarray) -> np.ndarray:
    """K = Z Z^T / M on pre-standardised Z (io.py:257-265)."""
    return (Z @ Z.T) / Z.shape[1]


def group_by_na(Y: np.ndarray, cols: list[int]) -> list[tuple[list[int], list[int]]]:
    """Group phenotype columns by NA pattern (io.py:187-203): (keep_rows, cols) per distinct pattern in
    first-seen order, keep_rows = strains observed in the group's columns."""
    mask = np.isnan(np.asarray(Y)[:, cols])
    groups: dict[bytes, list[int]] = {}
    for j, _ in enumerate(cols):
        groups.setdefault(np.packbits(mask[:, j]).tobytes(), []).append(j)
    out = []
    for js in groups.values():
        keep = np.flatnonzero(~mask[:, js[0]]).tolist()
        out.append((keep, [cols[j] for j in js]))
    return out


# --------------------------------------------------------------------------------------------------------
# plain GLS path (core.py:89-451)
# --------------------------------------------------------------------------------------------------------

def eigendecompose(K: np.ndarray, jitter: float = 1e-6):
    """eigh(K + jitter I), eigenvalues clamped at 0 (core.py:89-102)."""
    N = K.shape[0]
    s, U = np.linalg.eigh(K + jitter * np.eye(N))
    return np.clip(s, 0.0, None), U


def rotate(K, X, Y) -> dict:
    """Spectrum: s, U, X_rot = U^T X, Y_rot = U^T Y (core.py:105-116)."""
    s, U = eigendecompose(K)
    return {"s": s, "U": U, "X_rot": U.T @ X, "Y_rot": U.T @ Y}


def nll_eval(delta, s, X_rot, y_rot) -> float:
    """Profiled ML -2 log-likelihood at one delta for one column (core.py:122-145)."""
    N = X_rot.shape[0]
    Sd = s + delta
    w = 1.0 / Sd
    WX = w[:, None] * X_rot
    A = WX.T @ X_rot
    Xy = WX.T @ y_rot
    beta = np.linalg.solve(A, Xy)
    rWr = (y_rot * w * y_rot).sum() - (Xy * beta).sum()
    return N * (LOG2PI + np.log(rWr / N) + 1.0) + np.log(Sd).sum()


def _chol_batched(A):
    """Batched Cholesky with the reference's failure rule (core.py:166-170): a non-PD matrix is replaced by
    the identity factor and flagged."""
    n = A.shape[0]
    C = A.shape[-1]
    L = np.empty_like(A)
    bad = np.zeros(n, dtype=bool)
    for i in range(n):
        try:
            L[i] = np.linalg.cholesky(A[i])
        except np.linalg.LinAlgError:
            L[i] = np.eye(C)
            bad[i] = True
    return L, bad


def _chol_solve(L, B):
    """Solve (L L^T) X = B for a batch (cholesky_solve, core.py:173)."""
    out = np.empty_like(B)
    for i in range(L.shape[0]):
        z = np.linalg.solve(L[i], B[i])
        out[i] = np.linalg.solve(L[i].T, z)
    return out


def profile_loss(log_delta, s, X_rot, Y_rot, *, reml=False):
    """(G, P) -2 log-likelihood on a grid, ML or REML (core.py:148-187)."""
    log_delta = np.atleast_1d(np.asarray(log_delta, dtype=np.float64))
    N, C = X_rot.shape
    delta = np.exp(log_delta)
    Sd = s[None, :] + delta[:, None]  # (G, N)
    w = 1.0 / Sd
    WX = w[:, :, None] * X_rot[None, :, :]  # (G, N, C)
    A = np.einsum("gnc,nd->gcd", WX, X_rot)
    L, bad = _chol_batched(A)
    logdet_A = 2.0 * np.log(np.diagonal(L, axis1=-2, axis2=-1)).sum(-1)
    u = np.einsum("gnc,np->gcp", WX, Y_rot)  # (G, C, P)
    beta_x = _chol_solve(L, u)
    uAu = (u * beta_x).sum(axis=1)
    yWy = w @ (Y_rot * Y_rot)
    rWr = np.maximum(yWy - uAu, pos_floor())
    sum_log_Sd = np.log(Sd).sum(axis=1)
    if reml:
        loss = (N - C) * (LOG2PI + np.log(rWr / (N - C)) + 1.0) + sum_log_Sd[:, None] + logdet_A[:, None]
    else:
        loss = N * (LOG2PI + np.log(rWr / N) + 1.0) + sum_log_Sd[:, None]
    loss[bad, :] = np.inf
    return loss


def profile_loss_per_pheno(log_delta, s, X_rot, Y_rot, *, reml=False):
    """(P,) loss, one log delta per column (core.py:190-224)."""
    log_delta = np.asarray(log_delta, dtype=np.float64)
    N, C = X_rot.shape
    delta = np.exp(log_delta)
    Sd = s[None, :] + delta[:, None]  # (P, N)
    w = 1.0 / Sd
    WX = w[:, :, None] * X_rot[None, :, :]
    A = np.einsum("pnc,nd->pcd", WX, X_rot)
    L, bad = _chol_batched(A)
    logdet_A = 2.0 * np.log(np.diagonal(L, axis1=-2, axis2=-1)).sum(-1)
    u = np.einsum("pnc,np->pc", WX, Y_rot)
    beta_x = _chol_solve(L, u[:, :, None])[:, :, 0]
    uAu = (u * beta_x).sum(axis=1)
    yWy = (w * Y_rot.T * Y_rot.T).sum(axis=1)
    rWr = np.maximum(yWy - uAu, pos_floor())
    sum_log_Sd = np.log(Sd).sum(axis=1)
    if reml:
        loss = (N - C) * (LOG2PI + np.log(rWr / (N - C)) + 1.0) + sum_log_Sd + logdet_A
    else:
        loss = N * (LOG2PI + np.log(rWr / N) + 1.0) + sum_log_Sd
    loss[bad] = np.inf
    return loss


def _golden_refine(grid, loss_grid, per_pheno_loss, lo, hi, iters=50):
    """Grid argmin then vectorised golden section on [grid[i-1], grid[i+1]] (core.py:240-264,
    core.py:569-595).  Ties in argmin resolve to the first index, as torch.argmin does."""
    n_grid = grid.shape[0]
    idx = np.argmin(loss_grid, axis=0)
    lo_idx = np.clip(idx - 1, 0, n_grid - 1)
    hi_idx = np.clip(idx + 1, 0, n_grid - 1)
    a = grid[lo_idx].copy()
    b = grid[hi_idx].copy()
    c = a + INV_PHI2 * (b - a)
    d = a + INV_PHI * (b - a)
    fc = per_pheno_loss(c)
    fd = per_pheno_loss(d)
    for _ in range(iters):
        cond = fc < fd
        b = np.where(cond, d, b)
        a = np.where(cond, a, c)
        c = a + INV_PHI2 * (b - a)
        d = a + INV_PHI * (b - a)
        fc = per_pheno_loss(c)
        fd = per_pheno_loss(d)
    return np.clip(0.5 * (a + b), lo, hi)


def fit_delta_grid(spec, *, n_grid=64, log_delta_min=-10.0, log_delta_max=10.0, refine=True, reml=False):
    """Per-column log delta: grid then 50 golden-section iterations (core.py:227-264)."""
    s, X_rot, Y_rot = spec["s"], spec["X_rot"], spec["Y_rot"]
    grid = np.linspace(log_delta_min, log_delta_max, n_grid)
    loss = profile_loss(grid, s, X_rot, Y_rot, reml=reml)
    if not refine:
        return grid[np.argmin(loss, axis=0)]
    return _golden_refine(grid, loss, lambda ld: profile_loss_per_pheno(ld, s, X_rot, Y_rot, reml=reml),
                          log_delta_min, log_delta_max)


def _empty_result(M, n_real, P):
    z = lambda: np.zeros((M, n_real))  # noqa: E731
    return {"f": z(), "beta": z(), "se": z(), "sfve": z(), "nullh2": z(), "max_F": np.full(P, -np.inf)}


def snp_wald_scan(spec, log_delta, S_rot, *, n_real=None):
    """Plain GLS Wald scan with the C x C Schur complement (core.py:295-375).  Tiling in the reference is
    invisible in the results (chunking-invariance tests), so this restatement does one tile."""
    s, X_rot, Y_rot = spec["s"], spec["X_rot"], spec["Y_rot"]
    N, C = X_rot.shape
    P = Y_rot.shape[1]
    M = S_rot.shape[1]
    n_real = P if n_real is None else min(n_real, P)
    fl = pos_floor()
    delta = np.exp(log_delta)
    w = 1.0 / (s[None, :] + delta[:, None])  # (P, N)
    WX = w[:, :, None] * X_rot[None, :, :]
    A = np.einsum("pnc,nd->pcd", WX, X_rot)
    L, _bad = _chol_batched(A)
    u = np.einsum("pnc,np->pc", WX, Y_rot)
    beta_x = _chol_solve(L, u[:, :, None])[:, :, 0]
    r = Y_rot - X_rot @ beta_x.T  # (N, P)
    rWr = (w * r.T * r.T).sum(axis=1)
    wr = w * r.T
    B = np.einsum("pnc,nm->pcm", WX, S_rot)  # (P, C, M)
    AinvB = _chol_solve(L, B)
    sWs = w @ (S_rot * S_rot)
    quad = (B * AinvB).sum(axis=1)
    denom = np.maximum(sWs - quad, fl)
    num = wr @ S_rot
    beta = num / denom
    rss = np.maximum(rWr[:, None] - num * num / denom, fl)
    var_beta = rss / (N - C - 1) / denom
    f = beta * beta / var_beta
    out = _empty_result(M, n_real, P)
    out["max_F"] = f.max(axis=1) if M > 0 else np.full(P, -np.inf)
    out["f"] = f[:n_real].T.copy()
    out["beta"] = beta[:n_real].T.copy()
    out["se"] = np.sqrt(var_beta[:n_real]).T.copy()
    sfve = np.sqrt(np.maximum(num[:n_real] ** 2 / (denom[:n_real] * rWr[:n_real, None]), 0.0))
    out["sfve"] = sfve.T.copy()
    h2 = 1.0 / (1.0 + np.exp(log_delta[:n_real]))
    out["nullh2"] = np.broadcast_to(h2[None, :], (M, n_real)).copy()
    return out


def _loco(Z, X, Y, chrom, n_real, Z_kin, chrom_kin, per_chrom, on_chrom=None):
    """Shared chromosome loop of loco_scan / loco_scan_compat (core.py:403-436, core.py:694-726)."""
    if (Z_kin is None) != (chrom_kin is None):
        raise ValueError("Z_kin and chrom_kin go together")
    M = Z.shape[1]
    P = Y.shape[1]
    n_real = P if n_real is None else min(n_real, P)
    out = _empty_result(M, n_real, P)
    chrom = np.asarray(chrom)
    chroms = sorted(np.unique(chrom).tolist())
    for k, c in enumerate(chroms):
        test = chrom == c
        K = grm(Z[:, chrom != c]) if Z_kin is None else grm(Z_kin[:, np.asarray(chrom_kin) != c])
        res = per_chrom(K, Z[:, test])
        if n_real > 0:
            for key in ("f", "beta", "se", "sfve", "nullh2"):
                out[key][test, :] = res[key]
        out["max_F"] = np.maximum(out["max_F"], res["max_F"])
        if on_chrom is not None:
            on_chrom(k + 1, len(chroms))
    return out


def loco_scan(Z, X, Y, chrom, *, n_real=None, Z_kin=None, chrom_kin=None, n_grid=64, reml=False,
              on_chrom=None):
    """Plain-path LOCO scan (core.py:381-436)."""
    P = Y.shape[1]
    nr = P if n_real is None else min(n_real, P)

    def per_chrom(K, Z_test):
        spec = rotate(K, X, Y)
        ld = fit_delta_grid(spec, n_grid=n_grid, reml=reml)
        return snp_wald_scan(spec, ld, spec["U"].T @ Z_test, n_real=nr)
    return _loco(Z, X, Y, chrom, nr, Z_kin, chrom_kin, per_chrom, on_chrom)


def single_k_scan(Z, X, Y, *, n_real=None, Z_kin=None, n_grid=64, reml=False):
    """Plain-path single-kinship scan (core.py:439-451)."""
    spec = rotate(grm(Z if Z_kin is None else Z_kin), X, Y)
    ld = fit_delta_grid(spec, n_grid=n_grid, reml=reml)
    return snp_wald_scan(spec, ld, spec["U"].T @ Z, n_real=n_real)


# --------------------------------------------------------------------------------------------------------
# fastlmm-compat path (core.py:454-740)
# --------------------------------------------------------------------------------------------------------

def pinv(X):
    """Moore-Penrose pseudo-inverse with torch.linalg.pinv's default cut (rtol = eps * max(N, C)), the
    call at core.py:498."""
    return np.linalg.pinv(X, rtol=np.finfo(np.float64).eps * max(X.shape))


def compat_rotate(K, X, Y) -> dict:
    """Projected eigendecomposition (core.py:487-519): eigh((I - X X^+)(K + I)(I - X X^+)), drop the first
    D = ncols(X) eigenpairs, s = clamp(lambda - 1, 0), UY = U_eff^T (Y - X X^+ Y)."""
    N = K.shape[0]
    D = X.shape[1]
    Xpinv = pinv(X)
    Iperp = np.eye(N) - X @ Xpinv
    Kp = (Iperp @ (K + np.eye(N))) @ Iperp
    s_full, U_full = np.linalg.eigh(Kp)
    s = np.clip(s_full[D:N] - 1.0, 0.0, None)
    U_eff = np.ascontiguousarray(U_full[:, D:N])
    Y_r = Y - X @ (Xpinv @ Y)
    return {"s": s, "U_eff": U_eff, "X": X, "Xpinv": Xpinv, "UY": U_eff.T @ Y_r}


def profile_loss_compat(log_delta, s, UY):
    """(G, P) compat -2 log-likelihood (core.py:525-542)."""
    log_delta = np.atleast_1d(np.asarray(log_delta, dtype=np.float64))
    Neff = s.shape[0]
    Sd = s[None, :] + np.exp(log_delta)[:, None]
    w = 1.0 / Sd
    rWr = np.maximum(w @ (UY * UY), pos_floor())
    return Neff * (LOG2PI + np.log(rWr / Neff) + 1.0) + np.log(Sd).sum(axis=1)[:, None]


def profile_loss_compat_per_pheno(log_delta, s, UY):
    """(P,) compat loss, one log delta per column (core.py:545-558)."""
    Neff = s.shape[0]
    Sd = s[None, :] + np.exp(np.asarray(log_delta, dtype=np.float64))[:, None]
    w = 1.0 / Sd
    rWr = np.maximum((w * UY.T * UY.T).sum(axis=1), pos_floor())
    return Neff * (LOG2PI + np.log(rWr / Neff) + 1.0) + np.log(Sd).sum(axis=1)


def fit_delta_grid_compat(spec, *, n_grid=64, log_delta_min=-10.0, log_delta_max=10.0, refine=True):
    """Compat per-column log delta: grid then 50 golden-section iterations (core.py:561-595)."""
    s, UY = spec["s"], spec["UY"]
    grid = np.linspace(log_delta_min, log_delta_max, n_grid)
    loss = profile_loss_compat(grid, s, UY)
    if not refine:
        return grid[np.argmin(loss, axis=0)]
    return _golden_refine(grid, loss, lambda ld: profile_loss_compat_per_pheno(ld, s, UY),
                          log_delta_min, log_delta_max)


def compat_wald_from_sums(sWy, sWs, rWr, var_df):
    """Wald outputs from the two scan sums (core.py:649-667): returns f, beta, se, sfve (all FP64)."""
    fl = pos_floor()
    denom = np.maximum(sWs, fl)
    beta = sWy / denom
    rss = np.maximum(rWr - sWy * sWy / denom, fl)
    var_beta = rss / var_df / denom
    f = beta * beta / var_beta
    sfve = np.sqrt(np.maximum(sWy * sWy / (denom * rWr), 0.0))
    return f, beta, np.sqrt(var_beta), sfve


def snp_wald_scan_compat(spec, log_delta, S, *, n_real=None):
    """Compat Wald scan in the projected basis (core.py:601-671): S_r = S - X X^+ S, US = U_eff^T S_r, per
    (column, SNP) sWy = (w o UY)^T US, sWs = w^T (US o US), F with dof Neff - 1."""
    s, UY, X, Xpinv, U_eff = spec["s"], spec["UY"], spec["X"], spec["Xpinv"], spec["U_eff"]
    Neff = s.shape[0]
    P = UY.shape[1]
    M = S.shape[1]
    n_real = P if n_real is None else min(n_real, P)
    fl = pos_floor()
    US = U_eff.T @ (S - X @ (Xpinv @ S))
    w = 1.0 / (s[None, :] + np.exp(log_delta)[:, None])  # (P, Neff)
    rWr = np.maximum((w * UY.T * UY.T).sum(axis=1), fl)
    sWy = (w * UY.T) @ US  # (P, M)
    sWs = w @ (US * US)
    f, beta, se, sfve = compat_wald_from_sums(sWy, sWs, rWr[:, None], Neff - 1)
    out = _empty_result(M, n_real, P)
    out["max_F"] = f.max(axis=1) if M > 0 else np.full(P, -np.inf)
    out["f"] = f[:n_real].T.copy()
    out["beta"] = beta[:n_real].T.copy()
    out["se"] = se[:n_real].T.copy()
    out["sfve"] = sfve[:n_real].T.copy()
    h2 = 1.0 / (1.0 + np.exp(log_delta[:n_real]))
    out["nullh2"] = np.broadcast_to(h2[None, :], (M, n_real)).copy()
    return out


def loco_scan_compat(Z, X, Y, chrom, *, n_real=None, Z_kin=None, chrom_kin=None, n_grid=64, on_chrom=None,
                     return_log_delta=False):
    """Compat LOCO scan (core.py:677-726).  return_log_delta additionally returns {chrom: log_delta}."""
    P = Y.shape[1]
    nr = P if n_real is None else min(n_real, P)
    lds = {}

    def per_chrom(K, Z_test):
        spec = compat_rotate(K, X, Y)
        ld = fit_delta_grid_compat(spec, n_grid=n_grid)
        lds[len(lds)] = ld
        return snp_wald_scan_compat(spec, ld, Z_test, n_real=nr)
    out = _loco(Z, X, Y, chrom, nr, Z_kin, chrom_kin, per_chrom, on_chrom)
    return (out, lds) if return_log_delta else out


def single_k_scan_compat(Z, X, Y, *, n_real=None, Z_kin=None, n_grid=64):
    """Compat single-kinship scan (core.py:729-740)."""
    spec = compat_rotate(grm(Z if Z_kin is None else Z_kin), X, Y)
    ld = fit_delta_grid_compat(spec, n_grid=n_grid)
    return snp_wald_scan_compat(spec, ld, Z, n_real=n_real)
