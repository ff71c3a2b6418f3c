"""Pin the SSA replay statistically against the reference moment model.

The reference integrates the exact first and second moments of the same
linear network (scripts/model.jl:74-86); simulated cells must reproduce them
within sampling error for all five models, with and without capture.
"""
import numpy as np
import pytest

from abc_inference_transcription_amd.design import (MODEL_NAMES, REF_CHASE, REF_PULSE,
                                                    reference_design, CAPTURE_BETA)

AGES = [(2 * k + 1) / 10 * 20 for k in range(5)]


def z_scores(out, mom, cells_per_bc):
    """z of the per-(bin,cond) sample means and variances of u and l, with
    kurtosis-aware standard errors from the cells themselves."""
    u = out["cell_u"][0].astype(float)
    l = out["cell_l"][0].astype(float)
    b = out["cell_bin"][0]
    c = np.arange(len(u))
    zs = []
    for j in range(11):
        for k in range(5):
            sel = (b == k) & (c % 11 == j)
            for x, em, ev in ((u[sel], mom[j, k, 0], mom[j, k, 2]), (l[sel], mom[j, k, 1], mom[j, k, 4])):
                n = len(x)
                assert n == cells_per_bc
                if ev < 1e-6 or em * n < 20:
                    continue  # too few molecules in the sample for a normal z
                se_m = np.sqrt(ev / n)  # the model's variance sets the SE of the mean
                m4 = ((x - x.mean()) ** 4).mean()
                se_v = max(np.sqrt(max(m4 - x.var() ** 2, 0.0) / n), ev * np.sqrt(2.0 / n))
                zs.append((x.mean() - em) / se_m)
                zs.append((x.var(ddof=1) - ev) / se_v)
    return np.array(zs)


@pytest.mark.parametrize("model,row", [(1, 1), (2, 0), (3, 1), (4, 1), (5, 1)])
def test_ssa_matches_moments(oracle, golden_recovered, model, row):
    th_ref = golden_recovered[f"{MODEL_NAMES[model - 1]}__theta"][row]
    theta = np.append(th_ref, 0.5)[None, :]
    per = 160
    d = reference_design(cells=55 * per)
    od = oracle.OracleDesign(d)
    out = oracle.simulate(od, model, 7, 0, theta, n_threads=0)
    cfg = oracle.moment_cfg(AGES, REF_PULSE, REF_CHASE, 20.0, transient_mode=1)
    mom = oracle.recover(model, th_ref, cfg)
    z = z_scores(out, mom, per)
    assert abs(z.mean()) < 0.35, z.mean()
    assert np.abs(z).max() < 5.0
    assert 0.6 < z.std() < 1.4


@pytest.mark.parametrize("model,row", [(2, 3), (4, 2)])
def test_split_cells_match_moments(oracle, golden_recovered, model, row):
    """Cells split into many sub-lanes (shared gene path, 1/M of the births
    each) keep the law of the unsplit cell."""
    th_ref = golden_recovered[f"{MODEL_NAMES[model - 1]}__theta"][row]
    theta = np.append(th_ref, 0.5)[None, :]
    per = 160
    d = reference_design(cells=55 * per)
    d["split_target"], d["split_max"] = 20.0, 64
    od = oracle.OracleDesign(d)
    out = oracle.simulate(od, model, 9, 0, theta)
    cfg = oracle.moment_cfg(AGES, REF_PULSE, REF_CHASE, 20.0, transient_mode=1)
    mom = oracle.recover(model, th_ref, cfg)
    z = z_scores(out, mom, per)
    assert abs(z.mean()) < 0.35, z.mean()
    assert np.abs(z).max() < 5.0
    assert 0.6 < z.std() < 1.4


def test_ssa_capture_matches_downsample(oracle, golden_recovered):
    """Beta(2,8) capture thinning reproduces downsample (model.jl:221-239)
    with the Beta moments E=0.2, E[b^2]=6/110 (population variance)."""
    th_ref = golden_recovered["const_const__theta"][2]
    theta = np.append(th_ref, 0.0)[None, :]
    per = 200
    d = reference_design(cells=55 * per, capture_mode=CAPTURE_BETA)
    out = oracle.simulate(oracle.OracleDesign(d), 2, 3, 0, theta)
    cfg = oracle.moment_cfg(AGES, REF_PULSE, REF_CHASE, 20.0, transient_mode=1)
    mom = oracle.recover(2, th_ref, cfg)
    mb, m2b = 0.2, 6.0 / 110.0
    for j in range(11):
        mom[j] = oracle.downsample(mom[j], [mb] * 5, [m2b] * 5, [m2b - mb * mb] * 5)
    z = z_scores(out, mom, per)
    assert abs(z.mean()) < 0.35
    assert np.abs(z).max() < 5.0


def test_ssa_split_invariance(oracle):
    """Particle i uses stream first+i: a run split in two gives the same bits."""
    d = reference_design(cells=110)
    od = oracle.OracleDesign(d)
    th = oracle.sample_prior(od, 3, 5, 0, 4)
    a = oracle.simulate(od, 3, 5, 0, th)
    b = oracle.simulate(od, 3, 5, 2, th[2:])
    assert np.array_equal(a["cell_u"][2:], b["cell_u"])
    assert np.array_equal(a["stats_ref"][2:], b["stats_ref"])
    assert np.array_equal(a["steps"][2:], b["steps"])


def test_cell_stats_estimators(oracle):
    """The integer-exact statistics equal the data-side estimators of
    scripts/data_summary_statistics.jl computed with numpy (ddof=1)."""
    d = reference_design(cells=55 * 7)
    od = oracle.OracleDesign(d)
    rng = np.random.default_rng(0)
    C = d["cells"]
    u = rng.poisson(30, C).astype(np.uint32)
    l = rng.poisson(3, C).astype(np.uint32)
    c = np.arange(C)
    bins = ((c // 11) % 5).astype(np.uint8)
    sr, mo, ss = oracle.cell_stats(od, u, l, bins)
    w = np.asarray(d["weights"])
    for j in range(11):
        m1, m2, v1, v2, c12 = [], [], [], [], []
        for b in range(5):
            sel = (bins == b) & (c % 11 == j)
            x, y = u[sel].astype(float), l[sel].astype(float)
            m1.append(x.mean()); m2.append(y.mean()); v1.append(x.var(ddof=1)); v2.append(y.var(ddof=1))
            c12.append(np.cov(x, y)[0, 1])
            mm = mo[6 * (b + 5 * j):6 * (b + 5 * j) + 6]
            assert np.allclose(mm, [sel.sum(), x.mean(), y.mean(), x.var(ddof=1), np.cov(x, y)[0, 1], y.var(ddof=1)], rtol=1e-12)
            t = x + y
            assert np.isclose(ss[b + 5 * j], t.mean(), rtol=1e-14)
            assert np.isclose(ss[55 + b + 5 * j], t.var(ddof=1) / t.mean(), rtol=1e-12)
        sel = c % 11 == j
        assert np.isclose(sr[20 + j], l[sel].mean() / (u[sel].mean() + l[sel].mean()), rtol=1e-14)
        m1, m2, v1, v2, c12 = map(np.array, (m1, m2, v1, v2, c12))
        wj = w[:, j]
        wc = lambda a, b_: np.sum(wj * ((a - np.sum(wj * a)) * (b_ - np.sum(wj * b_))))
        tv1 = np.sum(wj * v1) + wc(m1, m1)
        tv2 = np.sum(wj * v2) + wc(m2, m2)
        assert np.isclose(sr[31 + j], np.sum(wj * c12) / np.sqrt(abs(tv1 * tv2)), rtol=1e-10)
        assert np.isclose(sr[42 + j], wc(m1, m2) / np.sqrt(abs(tv1 * tv2)), rtol=1e-10)
    for grp, conds in ((0, range(6)), (1, range(6, 11))):
        for b in range(5):
            sel = (bins == b) & np.isin(c % 11, list(conds))
            t = (u[sel] + l[sel]).astype(float)
            assert np.isclose(sr[2 * grp * 5 + b], t.mean(), rtol=1e-14)
            assert np.isclose(sr[(2 * grp + 1) * 5 + b], t.var(ddof=1) / t.mean(), rtol=1e-12)


def test_cell_stats_edge_cases(oracle):
    """Missing bins give zeros (the intent of get_mean_ccp_data's splice);
    a bin with one cell gives NaN variance like Julia's var; zero mean takes
    the 1e-4 guard."""
    d = reference_design(cells=22)
    od = oracle.OracleDesign(d)
    u = np.zeros(22, np.uint32)
    l = np.zeros(22, np.uint32)
    bins = np.zeros(22, np.uint8)
    bins[0] = 1  # cell 0 alone in bin 1 of condition 0
    sr, mo, ss = oracle.cell_stats(od, u, l, bins)
    assert mo[6 * 1 + 0] == 1 and np.isnan(mo[6 * 1 + 3])
    assert mo[6 * 2 + 0] == 0 and np.all(mo[6 * 2:6 * 3] == 0)
    assert sr[0] == 0.0 and sr[5] == 0.0          # all-zero counts: mean 0, FF 0/(0+1e-4)
    assert np.all(sr[20:31] == 0.0)               # ratio guard


@pytest.mark.parametrize("model,split", [(2, 0.0), (4, 0.0), (4, 2000.0)])
def test_fast_switching_gene_matches_moments(oracle, model, split):
    # Gene switching much faster than molecule events (most iterations are
    # switches, whose candidate molecule draws are discarded and redrawn);
    # with split cells every sub-lane also replays the fast gene path.
    kon, koff, al, ga = np.log10(20.0), np.log10(30.0), np.log10(4.0), np.log10(0.7)
    th_ref = [kon, koff, al, ga, np.log10(0.6)] if model == 2 else \
        [kon, koff] + [al + 0.1 * q for q in range(5)] + [ga, np.log10(0.6)]
    theta = np.append(th_ref, 0.3)[None, :]
    per = 200
    d = dict(reference_design(cells=55 * per), split_target=split)
    od = oracle.OracleDesign(d)
    out = oracle.simulate(od, model, 11, 0, theta, n_threads=0)
    cfg = oracle.moment_cfg(AGES, REF_PULSE, REF_CHASE, 20.0, transient_mode=1)
    mom = oracle.recover(model, th_ref, cfg)
    z = z_scores(out, mom, per)
    assert len(z) > 150
    assert abs(z.mean()) < 0.35, z.mean()
    assert np.abs(z).max() < 5.0
    assert 0.6 < z.std() < 1.4
