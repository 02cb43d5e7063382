"""Factored scan weights: 1 / z as a short positive exponential sum.

The scan's second product is sWs[m, p] = sum_k US[m, k]^2 / (s_k + delta_p) over all Neff eigenvalues
(reference fasterlmm/core.py:640-651, w = 1 / (s + delta) of core.py:640).  Writing

    1 / z = int_0^inf exp(-t z) dt = int_R exp(u) exp(-exp(u) z) du        (t = exp(u), z > 0)

and applying the trapezoidal rule in u with step h on [u_lo, u_hi] gives 1 / z ~= sum_r omega_r exp(-t_r z) with
t_r = exp(u_r), omega_r = h t_r > 0 (sinc quadrature; Stenger, "Numerical Methods Based on Sinc and Analytic
Functions", 1993).  The integrand is analytic in the strip |Im u| < pi / 2, so the discretisation error relative
to 1 / z is C exp(-pi^2 / h) for every z (C ~ 20 measured, hence h = pi^2 / log(20 / eps)); it decays like
exp(u) z as u -> -inf and double-exponentially as u -> +inf, so truncating at u_lo = log(eps / z_hi) - 1 and
u_hi = log(log(1 / eps) + 4) - log(z_lo) keeps both tails below eps on [z_lo, z_hi].  The term count grows like
log(z_hi / z_lo) / h: about 110 for z in [4.5e-5, 3e4] (log delta in [-10, 10] plus eigenvalues up to 8,000) at
eps = 1e-9.

With omega_r split evenly between the operands (lam_r = sqrt(omega_r)):

    sWs[m, p] = sum_r C[m, r] B[p, r],   C = (US o US) A,   A[k, r] = lam_r exp(-t_r s_k),
    B[p, r] = lam_r exp(-t_r delta_p),

a sum of positive terms over R ~ 100 entries instead of Neff, whose relative error is at most the sum's eps plus
the operands' rounding.  The even split keeps every row of C and of B within a bounded dynamic range, which the
row-scaled split-FP16 planes need; the exact weights 1 / (s_k + delta_p) span (s_max + delta) / (s_min + delta)
within one row, up to 1e8 for a singular kinship and a small delta.
"""
from __future__ import annotations

import math

import numpy as np

EPS = 1e-9  # relative accuracy of the sum on the scan's z range (the split operands round at ~1e-7)


def weight_nodes(z_lo: float, z_hi: float, eps: float = EPS) -> tuple[np.ndarray, np.ndarray]:
    """Nodes t (R,) and square-root weights lam (R,) of 1 / z ~= sum_r lam_r^2 exp(-t_r z), relative error <= ~eps
    for z in [z_lo, z_hi], 0 < z_lo <= z_hi < inf."""
    if not (0.0 < z_lo <= z_hi < math.inf) or not (0.0 < eps < 1e-2):
        raise ValueError(f"need 0 < z_lo <= z_hi < inf and 0 < eps < 0.01, got {z_lo!r}, {z_hi!r}, {eps!r}")
    h = math.pi ** 2 / math.log(20.0 / eps)
    u_lo = math.log(eps / z_hi) - 1.0
    u_hi = math.log(math.log(1.0 / eps) + 4.0) - math.log(z_lo)
    n = int(math.ceil((u_hi - u_lo) / h)) + 1
    t = np.exp(u_lo + h * np.arange(n, dtype=np.float64))
    return t, np.sqrt(h * t)


def z_range(s_min: float, s_max: float, log_delta_lo: float, log_delta_hi: float) -> tuple[float, float]:
    """The range of z = s_k + delta_p for eigenvalues in [s_min, s_max] and log delta in [lo, hi]."""
    return s_min + math.exp(log_delta_lo), s_max + math.exp(log_delta_hi)
