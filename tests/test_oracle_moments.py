"""Pin the moment-model restatement (scripts/model.jl) against the reference's
own outputs (data/recovered_statistics, golden slice) and closed forms."""
import math

import numpy as np
import pytest

from abc_inference_transcription_amd.design import MODEL_NAMES, REF_PULSE, REF_CHASE

AGES = [(2 * k + 1) / 10 * 20 for k in range(5)]


def ref_cfg(O, **kw):
    return O.moment_cfg(AGES, REF_PULSE, REF_CHASE, cycle=20.0, **kw)


@pytest.mark.parametrize("model", [1, 2, 3, 4, 5])
def test_recover_matches_reference_fixtures(oracle, golden_recovered, model):
    """recover_statistics.jl:27-68 outputs at the MAP sets, to CVODE tolerance
    (reltol 1e-3 in the reference; the gap is largest for the discontinuous
    decay-rate model)."""
    mn = MODEL_NAMES[model - 1]
    th = golden_recovered[f"{mn}__theta"][:4]
    want = golden_recovered[f"{mn}__stats"][:4]
    cfg = ref_cfg(oracle)
    rel_all = []
    for i in range(len(th)):
        got = oracle.recover(model, th[i], cfg)
        big = np.abs(want[i]) > 1e-3
        rel = np.abs(got - want[i])[big] / np.abs(want[i])[big]
        rel_all.append(rel)
        assert np.all(np.abs(got - want[i])[~big] < 5e-5)
    rel = np.concatenate(rel_all)
    assert np.median(rel) < 3e-3
    assert rel.max() < 0.06


def test_fixture_label_invariance(golden_recovered):
    """u + l and Var(u + l) do not depend on the labelling condition."""
    s = golden_recovered["const_const__stats"]  # rows, cond, age, 5
    tot = s[..., 0] + s[..., 1]
    assert np.max(np.abs(tot / tot[:, :1, :] - 1)) < 5e-3


def test_periodic_mean_closed_form(oracle):
    """Model 2 (constant, non-scaling): E[u+l](a) = m* + (m0 - m*) e^{-g a},
    m* = alpha p/gamma, m0 = m* (1 - e^{-g T}) / (2 - e^{-g T})."""
    th = np.array([0.3, -0.2, 1.1, -0.5, -0.1])
    kon, koff, al, ga = 10 ** th[:4]
    T = 20.0
    cfg = ref_cfg(oracle, transient_mode=1)
    out = oracle.recover(2, th, cfg)
    p = kon / (kon + koff)
    ms = al * p / ga
    m0 = ms * (1 - math.exp(-ga * T)) / (2 - math.exp(-ga * T))
    for a_i, a in enumerate(AGES):
        want = ms + (m0 - ms) * math.exp(-ga * a)
        for j in range(11):
            assert abs(out[j, a_i, 0] + out[j, a_i, 1] - want) / want < 1e-8


def test_transient_rule_vs_fixed_point(oracle):
    """The reference's 1% rule (model.jl:114-142) lands within a few percent of
    the exact periodic steady state."""
    th = np.array([-0.5, -1.2, 0.37, -0.65, -0.11])
    ss_ref, k = oracle.transient(2, th, ref_cfg(oracle, transient_mode=0))
    ss_fix, _ = oracle.transient(2, th, ref_cfg(oracle, transient_mode=1))
    assert 1 <= k <= 101
    idx = [0, 1, 3, 4, 6]
    assert np.max(np.abs(ss_ref[idx] - ss_fix[idx]) / np.abs(ss_fix[idx])) < 0.05


def test_periodic_boundary_is_binomial_halving(oracle):
    """Binomial(n, 1/2) of a fixed count: mean n/2, variance n/4
    (model.jl:98-111); covariances with the gene halve."""
    y = np.zeros(9)
    y[1], y[2] = 12.0, 4.0
    y[4], y[5] = 0.8, 0.2
    out = np.zeros(9)
    oracle.lib().orc_periodic_boundary(y.ctypes.data, out.ctypes.data)
    assert out[1] == 6.0 and out[2] == 2.0
    assert out[6] == 3.0 and out[8] == 1.0 and out[7] == 0.0
    assert out[4] == 0.4 and out[5] == 0.1
    # brute force over the binomial law for n = 7
    n = 7
    pk = np.array([math.comb(n, k) for k in range(n + 1)]) / 2 ** n
    ks = np.arange(n + 1)
    assert abs((pk * ks).sum() - n / 2) < 1e-12
    assert abs((pk * ks ** 2).sum() - (n / 2) ** 2 - n / 4) < 1e-12


def test_downsample_law_of_total_variance(oracle):
    """model.jl:235-237 is the Beta-Binomial law of total variance with the
    per-bin moments of beta; check it by exact enumeration on a tiny case."""
    betas = np.array([0.1, 0.25, 0.4])
    mb, m2b = betas.mean(), (betas ** 2).mean()
    vb_pop = betas.var()
    # counts: u ~ Poisson(3), l = 2 deterministic (independent)
    mu, vu, ml, vl, cv = 3.0, 3.0, 2.0, 0.0, 0.0
    s = np.array([[mu, ml, vu, cv, vl]])
    got = oracle.downsample(s, [mb], [m2b], [vb_pop])[0]
    # enumerate: U ~ Poisson(3) (truncated), L = 2, beta uniform over the list
    from scipy.stats import binom, poisson
    E_u = E_u2 = E_l = E_l2 = E_ul = 0.0
    for b in betas:
        for u in range(40):
            pu = poisson.pmf(u, 3.0)
            for x in range(u + 1):
                px = binom.pmf(x, u, b)
                for y in range(3):
                    py = binom.pmf(y, 2, b)
                    w = pu * px * py / len(betas)
                    E_u += w * x; E_u2 += w * x * x; E_l += w * y; E_l2 += w * y * y; E_ul += w * x * y
    assert abs(got[0] - E_u) < 1e-9 and abs(got[1] - E_l) < 1e-9
    assert abs(got[2] - (E_u2 - E_u ** 2)) < 1e-9
    assert abs(got[4] - (E_l2 - E_l ** 2)) < 1e-9
    assert abs(got[3] - (E_ul - E_u * E_l)) < 1e-9


def test_moment_stats_structure(oracle):
    """abc_simulation.jl:23-46 on hand-made moments: ratio, FF and
    correlation definitions."""
    J, nb = 11, 5
    s = np.zeros((J, nb, 5))
    s[..., 0] = 4.0
    s[..., 1] = 1.0
    s[..., 2] = 6.0
    s[..., 3] = 0.5
    s[..., 4] = 2.0
    w = np.full((nb, J), 0.2)
    out = oracle.moment_stats(s, w, 5, 6)
    assert np.allclose(out[0:5], 5.0) and np.allclose(out[10:15], 5.0)
    assert np.allclose(out[5:10], (6 + 1 + 2) / 5.0)
    assert np.allclose(out[20:31], 0.2)
    assert np.allclose(out[31:42], 0.5 / math.sqrt(6 * 2))
    assert np.allclose(out[42:53], 0.0)


def test_moment_sim_batch_is_the_composition(oracle):
    """The batched, threaded moment abc_sim (the CPU baseline of the GPU moment
    path) equals recover -> downsample per (age, pulse/chase) group ->
    moment_stats, particle by particle, bit for bit."""
    from abc_inference_transcription_amd.design import reference_design
    od = oracle.OracleDesign(reference_design(cells=55))
    th = oracle.sample_prior(od, 4, 0, 300, 3)
    is_chase = [0] * 6 + [1] * 5
    cfg = ref_cfg(oracle, iv=[0.0, 0.5] + [0.0] * 7)
    cap = np.random.default_rng(1).uniform(0.05, 0.3, (5, 2, 3))
    cap[:, :, 1] = cap[:, :, 0] ** 2 + cap[:, :, 2] * 0.1
    w = np.random.default_rng(2).uniform(0.1, 1.0, (5, 11))
    stats, mom = oracle.moment_sim(4, th, cfg, w, is_chase=is_chase, capture=cap.ravel(),
                                   n_threads=3)
    for i in range(3):
        s = oracle.recover(4, th[i], cfg)
        for j in range(11):
            g = is_chase[j]
            s[j] = oracle.downsample(s[j], cap[:, g, 0], cap[:, g, 1], cap[:, g, 2])
        assert np.array_equal(mom[i], s)
        assert np.array_equal(stats[i], oracle.moment_stats(s, w, 5, 6))


@pytest.mark.parametrize("model", [1, 2, 3, 4, 5])
def test_transient_affine_form_matches_direct_iteration(oracle, model):
    """The 1 % rule run on the cycle map's affine form (10 integrations) takes
    the same number of iterations as integrating every cycle (model.jl:128-141)
    and reaches the same state to integration accuracy."""
    from abc_inference_transcription_amd.design import reference_design
    od = oracle.OracleDesign(reference_design(cells=55))
    th = oracle.sample_prior(od, model, 0, 400 + model, 4)
    for iv in (None, [0.0, 0.5] + [0.0] * 7):
        cfg = ref_cfg(oracle, iv=iv)
        for i in range(len(th)):
            a, ka = oracle.transient(model, th[i], cfg)
            d, kd = oracle.transient(model, th[i], cfg, direct=True)
            assert ka == kd
            assert np.allclose(a, d, rtol=1e-7, atol=1e-12), (model, i, np.max(np.abs(a - d)))


@pytest.mark.parametrize("model", [1, 2, 3, 4, 5])
def test_whole_unlabelled_cycles_are_the_affine_map(oracle, model):
    """Trajectories that apply the affine cycle map to whole cycles away from
    the labelling window equal trajectories that integrate every cycle, to
    integration accuracy, at every condition and age."""
    from abc_inference_transcription_amd.design import reference_design
    od = oracle.OracleDesign(reference_design(cells=55))
    th = oracle.sample_prior(od, model, 0, 600 + model, 3)
    cfg = ref_cfg(oracle)
    for i in range(len(th)):
        a = oracle.recover(model, th[i], cfg)
        d = oracle.recover(model, th[i], cfg, direct=True)
        assert np.allclose(a, d, rtol=1e-7, atol=1e-12), (model, i, np.max(np.abs(a - d)))


def _solve(O, model, th, y0, method, rtol=1e-12, atol=1e-15, t1=20.0):
    import ctypes as C
    y = np.array(y0, np.float64)
    t = np.ascontiguousarray(th, np.float64)
    st = O.lib().orc_moment_solve(model, t.ctypes.data_as(C.c_void_p), 20.0, 5, 7.0, 2.0, 0.0,
                                  t1, y.ctypes.data_as(C.c_void_p), rtol, atol, method)
    return y, st


# log10 (kon, koff, alpha, gamma, lam) of the edge cases for the closed form:
# gamma = K (the 2K source resonates with nu = K + gamma), gamma a little
# above K (|d s| crosses 1/2 inside a piece), a stiff gene with a slow mRNA,
# gamma far above K, and every rate slow (nu s < 1/2: the series for E1(0)).
EGB_CASES = {
    "gamma=K": [0.0, 0.0, 1.0, math.log10(2.0), -0.3],
    "gamma~K": [0.0, 0.0, 1.0, math.log10(2.3), -0.3],
    "stiff gene": [3.0, 3.0, 1.0, -3.0, -0.3],
    "gamma>>K": [-3.0, -2.0, 2.0, math.log10(50.0), -0.1],
    "all slow": [-3.0, -3.0, 0.5, -3.0, -0.5],
}


@pytest.mark.parametrize("model", [1, 2])
@pytest.mark.parametrize("case", sorted(EGB_CASES))
def test_gene_block_closed_form_matches_integration(oracle, model, case):
    """Method 1 (the gene block E g, Var g, Cov gu, Cov gl in closed form, the
    rest by Dormand-Prince) against method 0 (Dormand-Prince on all nine
    moments) over one cycle through the labelling window, from a generic
    state, at rtol 1e-12: two different algorithms for the same linear ODE.
    Model 1 carries the size scaling alpha(t) = alpha (1 + t/T) (the al1
    terms), model 2 a constant alpha.  A dropped source term, a wrong decay
    rate or a sign error in any of the closed forms moves some moment by far
    more than the bound."""
    th = np.array(EGB_CASES[case] + [0.0])
    y0 = [0.3, 2.0, 0.5, 0.1, 0.05, 0.02, 3.0, 0.1, 0.6]
    a, _ = _solve(oracle, model, th, y0, 0)
    b, _ = _solve(oracle, model, th, y0, 1)
    assert np.all(np.isfinite(b))
    err = np.abs(b - a) / (np.abs(a) + 1e-9 * np.abs(a).max())
    assert err.max() < 1e-7, (case, err)


@pytest.mark.parametrize("model", [1, 2, 3, 4, 5])
def test_gene_block_closed_form_on_prior_draws(oracle, model):
    """The same comparison over reference-prior particles (kon, koff 1e-3..1e3,
    gamma 1e-3..1e2; per-bin rates for models 3-5), starting from the
    reference's transient start state; and method 1 takes fewer steps on
    the stiff ones.  Moments that have decayed to about zero carry method
    0's absolute error (atol 1e-15 per step over thousands of steps), hence
    the absolute floor of 1e-12."""
    from abc_inference_transcription_amd.design import reference_design
    od = oracle.OracleDesign(reference_design(cells=55))
    th = oracle.sample_prior(od, model, 0, 500 + model, 24)
    iv = [0.0, 0.5] + [0.0] * 7
    ratio = []
    for i in range(len(th)):
        a, s0 = _solve(oracle, model, th[i], iv, 0)
        b, s1 = _solve(oracle, model, th[i], iv, 1)
        assert np.all(np.abs(b - a) <= 1e-6 * np.abs(a) + 1e-12), (i, a, b)
        ratio.append(s0 / s1)
    assert np.median(ratio) > 1.0 and max(ratio) > 3.0
