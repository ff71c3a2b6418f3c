"""The Poisson sampler of the burn-in (inversion below mean 10, Hormann's
transformed rejection with squeeze above): its draws must follow the Poisson
pmf (chi-square goodness of fit over the central mass plus pooled tails),
with the right mean and variance, on both sides of the switch and far out."""
import numpy as np
import pytest
from scipy import stats as sst


@pytest.mark.parametrize("mu", [0.3, 2.5, 9.9, 10.0, 12.5, 37.5, 400.0, 2.5e4, 3.0e6])
def test_poisson_sampler_matches_pmf(oracle, mu):
    n = 200000
    x = oracle.poisson_sample(mu, 17, n).astype(np.int64)
    sd = np.sqrt(mu)
    assert abs(x.mean() - mu) < 5 * sd / np.sqrt(n)
    assert abs(x.var() / mu - 1.0) < 5 * np.sqrt(2.0 / n) + 1e-3
    lo, hi = int(max(0, np.floor(mu - 4 * sd))), int(np.ceil(mu + 4 * sd))
    if hi - lo > 200:  # pool into ~200 equal-probability-ish bins
        edges = np.unique(np.round(np.linspace(lo, hi + 1, 201)).astype(np.int64))
    else:
        edges = np.arange(lo, hi + 2)
    cdf = sst.poisson.cdf(edges - 1, mu)
    p = np.diff(cdf)
    p = np.concatenate([[cdf[0]], p, [1.0 - cdf[-1]]])
    obs = np.concatenate([[np.sum(x < edges[0])], np.histogram(x, bins=edges)[0],
                          [np.sum(x >= edges[-1])]])
    keep = p * n >= 5
    exp_ = p[keep] * n
    o = obs[keep].astype(float)
    o[-1] += obs[~keep].sum() if (~keep).any() else 0.0
    exp_[-1] += p[~keep].sum() * n if (~keep).any() else 0.0
    chi2 = ((o - exp_) ** 2 / exp_).sum()
    pval = sst.chi2.sf(chi2, len(o) - 1)
    assert pval > 1e-4, (mu, chi2, len(o), pval)


def test_poisson_sampler_edge_means(oracle):
    assert np.all(oracle.poisson_sample(0.0, 1, 100) == 0)
    assert np.all(oracle.poisson_sample(-1e-18, 1, 100) == 0)
    x = oracle.poisson_sample(1e-9, 1, 10000)
    assert x.sum() <= 2
