"""Model probabilities (scripts/model_probs.jl get_model_probs) in the
oracle's replay of the engine's bootstrap, pinned by what fixes them
independently: Julia's type-7 quantile against numpy's and Julia's known
values, the branch rules, the exact probabilities, and the law of the
resampled counts (binomial / multinomial, by chi-square); plus the
model_selection.jl decision rules in host.py on hand-built cases."""
import numpy as np
import pytest
from scipy import stats as sst

from abc_inference_transcription_amd import host


def test_julia_quantile(oracle):
    v = np.arange(1.0, 101.0)
    assert oracle.jl_quantile(v, 0.05) == pytest.approx(5.95, abs=1e-12)  # Julia: 5.95
    assert oracle.jl_quantile(v, 0.95) == pytest.approx(95.05, abs=1e-12)
    assert oracle.jl_quantile(np.array([3.5]), 0.3) == 3.5
    rng = np.random.default_rng(0)
    for n in (2, 7, 100, 256):
        x = np.sort(rng.normal(size=n))
        for p in (0.0, 0.05, 1.0 - 0.95, 0.5, 0.9, 0.95, 1.0):
            assert oracle.jl_quantile(x, p) == pytest.approx(np.quantile(x, p), rel=1e-14, abs=1e-15)


def test_branches_and_probabilities(oracle):
    counts = np.array([[0, 0, 0], [0, 7, 0], [3, 0, 0], [5, 5, 10], [1, 1, 0]])
    prob, lb, ub = oracle.model_probs(counts)
    assert np.all(prob[0] == 0) and np.all(lb[0] == 0) and np.all(ub[0] == 0)
    assert list(prob[1]) == [0, 1, 0] and list(lb[1]) == [0, 1, 0] and list(ub[1]) == [0, 1, 0]
    assert list(prob[2]) == [1, 0, 0] and list(ub[2]) == [1, 0, 0]
    assert list(prob[3]) == [0.25, 0.25, 0.5]
    assert list(prob[4]) == [0.5, 0.5, 0.0]
    assert np.all(lb[3:] <= ub[3:])
    assert np.all(lb[4, 2] == 0) and np.all(ub[4, 2] == 0)  # never drawn


def test_resampled_counts_are_binomial(oracle):
    """Two groups with l = (20, 30): each resample's count of group 1 is
    Binomial(50, 0.4); 400 genes x 256 resamples pooled."""
    counts = np.tile([20, 30], (400, 1))
    _, _, _, boot = oracle.model_probs(counts, seed=3, n_boot=256, boot=True)
    assert np.all(boot.sum(axis=2) == 50)
    x = boot[:, :, 0].ravel()
    k = np.arange(51)
    p = sst.binom.pmf(k, 50, 0.4)
    keep = p * len(x) >= 5
    obs = np.bincount(x, minlength=51)
    o = np.concatenate([obs[keep], [obs[~keep].sum()]])
    e = np.concatenate([p[keep], [p[~keep].sum()]]) * len(x)
    chi2 = ((o - e) ** 2 / e).sum()
    assert sst.chi2.sf(chi2, len(o) - 1) > 1e-4


def test_resampled_counts_tiny_and_multinomial(oracle):
    """S = 2 with l = (1, 1): 0, 1, 2 of group 1 with probabilities 1/4, 1/2,
    1/4; and three groups (1, 2, 7): the marginal of each group is
    Binomial(10, l / 10)."""
    _, _, _, boot = oracle.model_probs(np.tile([1, 1], (300, 1)), seed=5, n_boot=256, boot=True)
    obs = np.bincount(boot[:, :, 0].ravel(), minlength=3)
    e = np.array([0.25, 0.5, 0.25]) * obs.sum()
    assert sst.chi2.sf(((obs - e) ** 2 / e).sum(), 2) > 1e-4
    _, _, _, boot = oracle.model_probs(np.tile([1, 2, 7], (300, 1)), seed=6, n_boot=256, boot=True)
    for m, lm in enumerate((1, 2, 7)):
        x = boot[:, :, m].ravel()
        p = sst.binom.pmf(np.arange(11), 10, lm / 10)
        keep = p * len(x) >= 5
        ob = np.bincount(x, minlength=11)
        o = np.concatenate([ob[keep], [ob[~keep].sum()]])
        ex = np.concatenate([p[keep], [p[~keep].sum()]]) * len(x)
        assert sst.chi2.sf(((o - ex) ** 2 / ex).sum(), len(o) - 1) > 1e-4, m


def test_bounds_are_quantiles_of_the_resamples(oracle):
    counts = np.array([[40, 60], [3, 9]])
    prob, lb, ub, boot = oracle.model_probs(counts, seed=9, n_boot=100, boot=True)
    for g in range(2):
        S = counts[g].sum()
        for m in range(2):
            sh = np.sort(boot[g, :, m] / S)
            assert lb[g, m] == oracle.jl_quantile(sh, 1.0 - 0.95)
            assert ub[g, m] == oracle.jl_quantile(sh, 0.95)


def test_selection_rules():
    # genes 1..5; two columns (constant, non-constant)
    prob = np.array([[0.9, 0.1], [0.2, 0.8], [0.5, 0.5], [0.0, 0.0], [0.7, 0.3]])
    lb = np.array([[0.8, 0.05], [0.1, 0.7], [0.3, 0.3], [0.0, 0.0], [0.5, 0.2]])
    ub = np.array([[0.95, 0.2], [0.3, 0.9], [0.7, 0.7], [0.0, 0.0], [0.8, 0.5]])
    n = np.array([10, 10, 10, 10, 2])
    thr = (2 / 3) / (1 + 2 / 3)  # 0.4, model_selection.jl:103-104
    c, nc, und, nf = host.get_c_nc_genes(prob, lb, ub, [1, 2, 3, 4, 5], n, 5, thr)
    assert list(c) == [1] and list(nc) == [2] and list(und) == [3] and list(nf) == [4, 5]
    s1, s2, s0 = host.cluster_c_genes(prob, lb, ub, [1, 2, 3, 4])
    assert list(s1) == [1] and list(s2) == [2] and list(s0) == [3, 4]


def test_nc_cluster_rules():
    # three models; gene 1: model 1 beats both; gene 2: 1 beats 2 and 3 beats 2
    # (two-model); gene 3: only 2 beats 3 (maybe 2); gene 4: nothing; gene 5:
    # every model beats one other (all three: no model with M - 1 wins)
    prob = np.full((5, 3), 1.0 / 3)
    lb = np.array([[0.6, 0.1, 0.1], [0.4, 0.1, 0.4], [0.2, 0.4, 0.1], [0.2, 0.2, 0.2],
                   [0.5, 0.5, 0.5]])
    ub = np.array([[0.8, 0.3, 0.3], [0.5, 0.3, 0.5], [0.5, 0.6, 0.3], [0.5, 0.5, 0.5],
                   [0.5, 0.9, 0.9]])
    sel, (one, two, maybe, und) = host.cluster_nc_genes(prob, lb, ub, [1, 2, 3, 4, 5])
    assert sel[0] == [1] and sel[1] == [1, 3] and sel[2] == [-2] and sel[3] == [0]
    assert list(one[0]) == [1] and list(two[1]) == [2] and list(maybe[1]) == [3]
    assert 4 in list(und[0])
    # gene 5: model 1 beats 2 and 3 (lb 0.5 >= ub 0.9? no) -> only lb >= ub pairs:
    # 1 vs 2: 0.5 >= 0.9 no; 1 vs 3 no; 2 vs 1: 0.5 >= 0.5 yes; 2 vs 3 no;
    # 3 vs 1 yes; 3 vs 2 no -> nc = [2, 3]: two models
    assert sel[4] == [2, 3] and 5 in list(two[2])
