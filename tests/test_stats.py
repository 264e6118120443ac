"""Film statistics used by the parity and equal-time tests (lm_amd.stats):
the Welch test holds its level on identical distributions and detects a
shifted mean, and the reference-noise-corrected rMSE is unbiased where the
raw rMSE is inflated by the reference's own variance."""
import numpy as np
import pytest

from lm_amd import stats


def sums(rng, mu, sigma, n, shape):
    x = rng.normal(mu, sigma, size=(n,) + shape)
    return x.sum(0), (x * x).sum(0)


def test_welch_level_and_power():
    rng = np.random.default_rng(1)
    shape = (64, 64, 3)
    mu = rng.uniform(0.5, 2.0, size=shape)
    sa, qa = sums(rng, mu, 1.0, 64, shape)
    sb, qb = sums(rng, mu, 1.0, 32, shape)
    frac, _, a = stats.welch(sa, qa, 64, sb, qb, 32, channel_reduce='any')
    assert frac <= stats.binomial_bound(a, 64 * 64)
    sc, qc = sums(rng, mu + 1.0, 1.0, 32, shape)
    frac2, _, _ = stats.welch(sa, qa, 64, sc, qc, 32, channel_reduce='any')
    assert frac2 > 0.99


@pytest.mark.parametrize('regime', ['dark', 'bright'])
def test_rel_mse_corrected_removes_reference_noise(regime):
    """Dark pixels (R^2 << eps: constant weight) and bright pixels with a
    reference of small relative noise: the corrected estimate matches the
    error against the true image where the raw one is inflated by Var(R)."""
    rng = np.random.default_rng(2 if regime == 'dark' else 3)
    shape = (96, 96, 3)
    if regime == 'dark':
        mu, sigma = rng.uniform(0.0, 0.03, size=shape), rng.uniform(0.05, 0.2, size=shape)
    else:
        mu, sigma = rng.uniform(1.0, 2.0, size=shape), rng.uniform(0.5, 2.0, size=shape)
    n_img, n_ref = 64, 256
    truth = np.mean(sigma ** 2 / n_img / (mu ** 2 + 1e-2))   # E rMSE of the image against mu
    raw, cor = [], []
    for _ in range(8):
        si, _ = sums(rng, mu, sigma, n_img, shape)
        sr, qr = sums(rng, mu, sigma, n_ref, shape)
        img = si / n_img
        raw.append(stats.rel_mse(img, sr / n_ref))
        cor.append(stats.rel_mse_corrected(img, sr, qr, n_ref))
    raw, cor = np.mean(raw), np.mean(cor)
    print(regime, truth, raw, cor)
    # the raw error carries the reference's variance, sigma^2 / n_ref (a quarter of the image's)
    assert raw > 1.15 * truth
    assert abs(cor / truth - 1) < 0.05, (cor, truth, raw)
