"""Prior draws of the oracle (the engine's must match them bit for bit).

The reference draws with Julia's unseeded global RNG (abc_simulation.jl:3-11),
so no individual draw can be matched; these tests pin the law: ranges, column
order (get_vary_map, model.jl:30-43), uniformity, independence, the per-bin
fold-change construction of the configured priors, and that a particle's
draw depends only on its index."""
import numpy as np
import pytest
from scipy import stats as sst

from abc_inference_transcription_amd.design import config_design, reference_design, n_params

N = 20000
FC = np.log10(5.0)


def cols(model, nb=5):
    """Column ranges of theta: kon..., koff, alpha..., gamma..., lambda, splice."""
    nk = nb if model == 3 else 1
    na = nb if model == 4 else 1
    ng = nb if model == 5 else 1
    o = np.cumsum([0, nk, 1, na, ng, 1, 1])
    return {k: slice(o[i], o[i + 1]) for i, k in enumerate(["kon", "koff", "al", "ga", "lam", "spl"])}


def ks_uniform(x, lo, hi):
    return sst.kstest((x - lo) / (hi - lo), "uniform").pvalue


@pytest.mark.parametrize("model", [1, 2, 3, 4, 5])
def test_reference_prior_law(oracle, model):
    od = oracle.OracleDesign(reference_design(cells=10))
    th = oracle.sample_prior(od, model, 0, 0, N)
    assert th.shape == (N, n_params(model)) == (N, (5 if model < 3 else 9) + 1)
    c = cols(model)
    ranges = {"kon": (-3, 3), "koff": (-3, 3), "al": (-3, 3), "ga": (-3, 2), "lam": (-0.7, 0)}
    for k, (lo, hi) in ranges.items():
        x = th[:, c[k]]
        assert x.min() >= lo and x.max() < hi, k
        for j in range(x.shape[1]):   # every varying bin is its own uniform draw
            assert ks_uniform(x[:, j], lo, hi) > 1e-3, (k, j)
    assert np.all(th[:, c["spl"]] == 0.0)    # splicing column fixed at the design value
    r = np.corrcoef(th[:, :-1].T)
    assert np.max(np.abs(r - np.eye(r.shape[0]))) < 0.04


@pytest.mark.parametrize("model", [1, 2, 3, 4, 5])
def test_config_prior_law(oracle, model):
    od = oracle.OracleDesign(config_design(cells=10))
    th = oracle.sample_prior(od, model, 0, 0, N)
    c = cols(model)
    kon, koff, al, ga = th[:, c["kon"]], th[:, c["koff"]][:, 0], th[:, c["al"]], th[:, c["ga"]]
    size = al - koff[:, None]               # log10 burst size = alpha / koff
    base = {"kon": (kon, -2, 2), "size": (size, -1, 3), "ga": (ga, -2, 1)}
    varying = {3: "kon", 4: "size", 5: "ga"}.get(model)
    for k, (x, lo, hi) in base.items():
        if k == varying:
            # base + per-bin log-uniform fold change in [1/5, 5]: bins differ by
            # a difference of two U(-FC, FC), a triangular law on (-2FC, 2FC)
            assert x.shape[1] == 5
            assert x.min() >= lo - FC and x.max() < hi + FC
            d = x[:, 0] - x[:, 1]
            assert np.abs(d).max() < 2 * FC
            p = sst.kstest(d, sst.triang(c=0.5, loc=-2 * FC, scale=4 * FC).cdf).pvalue
            assert p > 1e-3
            # shared base: corr = Var(base) / (Var(base) + Var(fold change))
            w2, f2 = (hi - lo) ** 2, (2 * FC) ** 2
            assert abs(np.corrcoef(x[:, 0], x[:, 1])[0, 1] - w2 / (w2 + f2)) < 0.01
        else:
            assert x.shape[1] == 1
            assert x.min() >= lo and x.max() < hi, k
            assert ks_uniform(x[:, 0], lo, hi) > 1e-3, k
    assert koff.min() >= -3 and koff.max() < 3 and ks_uniform(koff, -3, 3) > 1e-3
    lam, spl = th[:, c["lam"]][:, 0], th[:, c["spl"]][:, 0]
    assert ks_uniform(lam, -0.7, 0.0) > 1e-3 and ks_uniform(spl, -1.0, 2.0) > 1e-3


def test_prior_depends_only_on_particle_index(oracle):
    od = oracle.OracleDesign(config_design(cells=10))
    a = oracle.sample_prior(od, 3, 0, 0, 100)
    b = oracle.sample_prior(od, 3, 0, 37, 20)
    np.testing.assert_array_equal(a[37:57], b)
    c = oracle.sample_prior(od, 3, 1, 0, 100)           # another seed: other draws
    assert not np.any(a == c)
    d = oracle.sample_prior(od, 4, 0, 0, 100)           # another model: other stream
    assert not np.any(a[:, 6] == d[:, 6])


def test_cell_streams_differ_between_models(oracle):
    """The model is part of every cell's Philox counter (cell | model << 24), so
    the same particle index and cell draw different set-ups under different
    models: models 1 and 2 (the same parameter columns, so the same rates
    here) draw each cell's capture efficiency from their own stream, and the
    same model twice draws identical ones."""
    from abc_inference_transcription_amd.design import reference_design, CAPTURE_BETA
    od = oracle.OracleDesign(reference_design(cells=64, capture_mode=CAPTURE_BETA))
    th = np.array([0.5, 0.3, 0.8, -0.2, -0.3, 0.0])
    a = oracle.burn_cells(od, 1, 0, 5, th, 64)
    b = oracle.burn_cells(od, 2, 0, 5, th, 64)
    c = oracle.burn_cells(od, 1, 0, 5, th, 64)
    np.testing.assert_array_equal(a["capture"], c["capture"])
    assert np.mean(a["capture"] != b["capture"]) > 0.95
