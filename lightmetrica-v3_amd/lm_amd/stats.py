"""Film statistics: per-pixel Welch t-test between two renders and relative MSE.

Inputs are the per-pixel sums and sums of squares of the per-sample
estimates (FP64) that renderer::pt_mi355x and the oracle both produce.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import stats as _st


def mean_var(s: np.ndarray, q: np.ndarray, n: int):
    """Sample mean and unbiased variance from sum / sum of squares."""
    m = s / n
    v = np.maximum(q / n - m * m, 0.0) * (n / max(n - 1, 1))
    return m, v


def welch(sa, qa, na, sb, qb, nb, alpha: float = 0.01, channel_reduce: str = 'luminance'):
    """Per-pixel two-sided Welch t-test of equal means.

    With channel_reduce='luminance' the test runs on Y = 0.2126 R + 0.7152 G
    + 0.0722 B, whose per-sample sum is linear in the channel sums but whose
    sum of squares is not; the test then uses the mean of the per-channel
    variances weighted by the squared coefficients as a conservative bound
    (Var(Y) <= (sum_c w_c sigma_c)^2 by Cauchy-Schwarz). With
    channel_reduce='any', a pixel fails if any channel fails.
    Returns (fail_fraction, p_values, expected_fraction_under_H0)."""
    ma, va = mean_var(sa, qa, na)
    mb, vb = mean_var(sb, qb, nb)
    if channel_reduce == 'luminance':
        w = np.array([0.2126, 0.7152, 0.0722])
        ma_ = (ma * w).sum(-1)
        mb_ = (mb * w).sum(-1)
        va_ = ((np.sqrt(va) * w).sum(-1)) ** 2
        vb_ = ((np.sqrt(vb) * w).sum(-1)) ** 2
        p = _pvalue(ma_, va_, na, mb_, vb_, nb)
        return float((p < alpha).mean()), p, alpha
    p = _pvalue(ma, va, na, mb, vb, nb)
    fail = (p < alpha).any(-1)
    return float(fail.mean()), p, 1 - (1 - alpha) ** 3


def _pvalue(ma, va, na, mb, vb, nb):
    se2 = va / na + vb / nb
    diff = ma - mb
    with np.errstate(divide='ignore', invalid='ignore'):
        t = diff / np.sqrt(se2)
        dof = se2 ** 2 / ((va / na) ** 2 / max(na - 1, 1) + (vb / nb) ** 2 / max(nb - 1, 1))
    p = np.ones_like(diff)
    ok = se2 > 0
    p[ok] = 2 * _st.t.sf(np.abs(t[ok]), np.maximum(dof[ok], 1.0))
    # both estimates deterministic: fail only when the means differ
    det = ~ok
    p[det] = np.where(np.abs(diff[det]) <= 1e-9 * (1 + np.abs(ma[det])), 1.0, 0.0)
    return p


def binomial_bound(alpha: float, n: int, sigmas: float = 4.0) -> float:
    """Upper bound on the failing fraction expected under H0 (alpha +
    `sigmas` binomial standard deviations over n independent tests)."""
    return alpha + sigmas * math.sqrt(alpha * (1 - alpha) / max(n, 1))


def rel_mse(img: np.ndarray, ref: np.ndarray, eps: float = 1e-2) -> float:
    """rMSE = mean((I - R)^2 / (R^2 + eps)) over pixels and channels."""
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))


def rel_mse_corrected(img: np.ndarray, ref_sum: np.ndarray, ref_sumsq: np.ndarray, n_ref: int,
                      eps: float = 1e-2) -> float:
    """rMSE of `img` against the true image, estimated from a finite
    reference R (per-pixel sum and sum of squares of n_ref independent
    samples): for I independent of R with E[R] = mu,
    E[(I - R)^2] = E[(I - mu)^2] + Var(R), and s^2 / n_ref (the unbiased
    sample variance over n_ref) estimates Var(R) without bias, so
    mean(((I - R)^2 - s^2 / n_ref) / (R^2 + eps)) removes the reference's own
    noise from the error. The weight 1 / (R^2 + eps) uses the noisy R, which
    leaves a relative bias of about 3 Var(R) / R^2 in pixels with R^2 >> eps
    (none where eps dominates): small for a reference whose own relative
    noise is small."""
    ref, var = mean_var(ref_sum, ref_sumsq, n_ref)
    return float(np.mean(((img - ref) ** 2 - var / n_ref) / (ref ** 2 + eps)))
