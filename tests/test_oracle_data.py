"""The data-side summary statistics and their bootstrap SEs
(scripts/data_summary_statistics.jl) in the CPU restatement, against a
literal numpy restatement of the reference's functions: point estimates on
synthetic cells with every edge case the estimators have (missing cluster,
single-cell cluster, zero mean, empty condition, zero covariance), and the
standard errors replayed from the same resample positions."""
import numpy as np
import pytest
from scipy import stats as sps

NC, J = 5, 11


# ---- the reference's functions, restated with numpy (Julia semantics) ----
def jl_var(x):
    return np.nan if len(x) == 1 else float(np.sum((x - x.mean()) ** 2) / (len(x) - 1))


def jl_cov(x, y):
    return np.nan if len(x) == 1 else float(np.sum((x - x.mean()) * (y - y.mean())) / (len(x) - 1))


def ccp(data, ages, nc):
    """get_mean_ccp_data / get_ff_ccp_data (data_summary_statistics.jl:2-32);
    a missing cluster gives 0 at its own index."""
    m, ff = np.zeros(nc), np.zeros(nc)
    for c in range(nc):
        x = data[ages == c].astype(float)
        if len(x) == 0:
            continue
        m[c] = x.mean()
        eps = 0.0001 if m[c] == 0 else 0.0
        ff[c] = jl_var(x) / (m[c] + eps)
    return m, ff


def ratio(u, l):
    """get_cond_specific_ratios (data_summary_statistics.jl:71-77)"""
    if len(u) == 0:
        return 0.0
    mu, ml = u.mean(), l.mean()
    return ml / (mu + ml) if mu + ml > 0 else 0.0


def wcov(x, y, w):
    """weighted_cov (data_summary_statistics.jl:179-181)"""
    return float(np.sum(w * ((x - np.sum(w * x)) * (y - np.sum(w * y)))))


def corr(u, l, ages, nc, w):
    """get_cond_specific_correlations (data_summary_statistics.jl:111-147)"""
    m1, m2, v1, v2, c12 = (np.zeros(nc) for _ in range(5))
    for c in range(nc):
        sel = ages == c
        if not sel.any():
            continue
        x, y = u[sel].astype(float), l[sel].astype(float)
        m1[c], m2[c] = x.mean(), y.mean()
        v1[c], v2[c], c12[c] = jl_var(x), jl_var(y), jl_cov(x, y)
    tv1 = np.sum(w * v1) + wcov(m1, m1, w)
    tv2 = np.sum(w * v2) + wcov(m2, m2, w)
    nz = tv1 != 0 and tv2 != 0
    mc = np.sum(w * c12) / np.sqrt(abs(tv1 * tv2)) if (np.any(c12 != 0) and nz) else 0.0
    cm = wcov(m1, m2, w) / np.sqrt(abs(tv1 * tv2)) if nz else 0.0
    return mc, cm


def one_set(u, l, age, cond, pulse, chase, w, nc, J):
    """the 4 nc + 3 J statistics of get_summary_stats for one gene"""
    pm, pf = ccp(u[pulse] + l[pulse], age[pulse], nc)
    cm_, cf = ccp(u[chase] + l[chase], age[chase], nc)
    r = [ratio(u[cond == j], l[cond == j]) for j in range(J)]
    mc, cmn = zip(*[corr(u[cond == j], l[cond == j], age[cond == j], nc, w[:, j]) for j in range(J)])
    return np.concatenate([pm, pf, cm_, cf, r, mc, cmn])


def boot_weights(age, cond, nc, J):
    """get_correlation_se's weights: each condition's cluster shares in the resample"""
    w = np.zeros((nc, J))
    for j in range(J):
        tot = np.sum(cond == j)
        for c in range(nc):
            w[c, j] = np.sum((cond == j) & (age == c)) / tot if tot else 0.0
    return w


def to_columns(pt, se, nc, J):
    """the oracle's column order: the 14 summary files"""
    cols = []
    for f in range(2):
        cols += [pt[2 * f * nc:(2 * f + 1) * nc], pt[(2 * f + 1) * nc:(2 * f + 2) * nc],
                 se[2 * f * nc:(2 * f + 1) * nc], se[(2 * f + 1) * nc:(2 * f + 2) * nc]]
    for q in range(3):
        cols += [pt[4 * nc + q * J:4 * nc + (q + 1) * J], se[4 * nc + q * J:4 * nc + (q + 1) * J]]
    return np.concatenate(cols)


def synthetic(seed, n=400, G=4, nc=NC, J=J):
    rng = np.random.default_rng(seed)
    age = rng.integers(0, nc, n)
    cond = rng.integers(-1, J, n)
    u = rng.poisson(rng.uniform(0.5, 6, G), (n, G))
    l = rng.poisson(rng.uniform(0.1, 3, G), (n, G))
    pulse = np.nonzero((cond >= 0) & (cond < 6))[0]
    chase = np.nonzero(cond >= 6)[0]
    w = rng.random((nc, J))
    w /= w.sum(axis=0)
    return u, l, age, cond, pulse, chase, w


def edge_case():
    """cells chosen so that every estimator branch is taken: cluster 4 has no
    pulse cell, cluster 3 of the pulse cells has one cell, cluster 4 of the
    chase cells has only zero counts, condition 10 has no cell, condition 9
    has zero labelled counts, condition 8 has one cell per cluster (NaN
    variances), condition 7 has u constant within clusters (zero cov)."""
    rng = np.random.default_rng(7)
    cells = []
    for j in range(6):                       # pulse conditions
        for c in range(3):
            for _ in range(6):
                cells.append((c, j, rng.integers(0, 9), rng.integers(0, 4)))
    cells.append((3, 2, 5, 1))               # the single pulse cell of cluster 3
    def chase_cell(c, j, uu, ll):            # chase cluster 4 holds zeros only
        cells.append((c, j, 0 if c == 4 else uu, 0 if c == 4 else ll))
    for c in range(5):                       # condition 6: chase
        for _ in range(5):
            chase_cell(c, 6, rng.integers(1, 9), rng.integers(0, 4))
    for c in range(5):                       # condition 7: u constant per cluster
        for _ in range(4):
            chase_cell(c, 7, c + 1, rng.integers(0, 5))
    for c in range(5):                       # condition 8: one cell per cluster
        chase_cell(c, 8, rng.integers(0, 6), rng.integers(0, 3))
    for c in range(5):                       # condition 9: no labelled molecule
        for _ in range(3):
            chase_cell(c, 9, rng.integers(0, 6), 0)
    for _ in range(4):                       # cells in no condition
        cells.append((0, -1, 3, 3))
    a = np.array(cells)
    age, cond, u, l = a[:, 0], a[:, 1], a[:, 2:3], a[:, 3:4]
    pulse = np.nonzero((cond >= 0) & (cond < 6))[0]
    chase = np.nonzero(cond >= 6)[0]
    w = np.full((NC, J), 0.2)
    return u, l, age, cond, pulse, chase, w


@pytest.mark.parametrize("case", ["synthetic", "edge"])
def test_point_estimates_match_reference_functions(oracle, case):
    u, l, age, cond, pulse, chase, w = synthetic(3) if case == "synthetic" else edge_case()
    out = oracle.data_stats(u, l, age, cond, NC, J, pulse, chase, w, n_boot=3)
    for g in range(u.shape[1]):
        want = one_set(u[:, g], l[:, g], age, cond, pulse, chase, w, NC, J)
        got_pt = np.concatenate([out[g, 0:2 * NC], out[g, 4 * NC:6 * NC]] +
                                [out[g, 8 * NC + 2 * q * J:8 * NC + (2 * q + 1) * J] for q in range(3)])
        np.testing.assert_allclose(got_pt, want, rtol=1e-12, atol=1e-15, equal_nan=True)
    if case == "edge":
        o = out[0]
        assert o[4] == 0.0 and o[NC + 4] == 0.0            # missing pulse cluster
        assert np.isnan(o[NC + 3])                          # single-cell cluster: var NaN
        assert o[4 * NC + 4] == 0.0 and o[5 * NC + 4] == 0.0  # chase cluster of zeros: mean 0, FF 0
        r = o[8 * NC:8 * NC + J]
        assert r[10] == 0.0 and r[9] == 0.0                # empty condition, no labelled molecules
        mc = o[8 * NC + 2 * J:8 * NC + 3 * J]
        cm = o[8 * NC + 4 * J:8 * NC + 5 * J]
        assert mc[10] == 0.0 and cm[10] == 0.0
        assert np.isnan(mc[8]) and np.isnan(cm[8])         # one cell per cluster: NaN moments
        assert mc[7] == 0.0                                 # every covariance zero


def test_bootstrap_se_replays_the_resamples(oracle):
    """Each SE is the n-1 standard deviation over B resamples, where resample b
    draws four independent samples with replacement (pulse cells, chase cells,
    all cells for the ratios, all cells for the correlations, the latter
    weighted by its own cluster shares per condition)."""
    u, l, age, cond, pulse, chase, w = synthetic(11, n=160, G=2)
    B, seed = 12, 5
    out, bt = oracle.data_stats(u, l, age, cond, NC, J, pulse, chase, w, n_boot=B, seed=seed,
                                boot=True)
    n = len(age)
    for g in range(u.shape[1]):
        sets = []
        for b in range(B):
            ip = pulse[oracle.data_resample(g, b, 0, len(pulse), seed)]
            ic = chase[oracle.data_resample(g, b, 1, len(chase), seed)]
            ir = oracle.data_resample(g, b, 2, n, seed)
            ik = oracle.data_resample(g, b, 3, n, seed)
            ug, lg = u[:, g], l[:, g]
            pm, pf = ccp(ug[ip] + lg[ip], age[ip], NC)
            cmn, cf = ccp(ug[ic] + lg[ic], age[ic], NC)
            r = [ratio(ug[ir][cond[ir] == j], lg[ir][cond[ir] == j]) for j in range(J)]
            wb = boot_weights(age[ik], cond[ik], NC, J)
            cc = [corr(ug[ik][cond[ik] == j], lg[ik][cond[ik] == j], age[ik][cond[ik] == j], NC,
                       wb[:, j]) for j in range(J)]
            mc, cm = zip(*cc)
            sets.append(np.concatenate([pm, pf, cmn, cf, r, mc, cm]))
        sets = np.array(sets)
        np.testing.assert_allclose(bt[g], sets, rtol=1e-12, atol=1e-15, equal_nan=True)
        se = np.sqrt(np.sum((sets - sets.mean(axis=0)) ** 2, axis=0) / (B - 1))
        pt = one_set(u[:, g], l[:, g], age, cond, pulse, chase, w, NC, J)
        np.testing.assert_allclose(out[g], to_columns(pt, se, NC, J), rtol=1e-10, atol=1e-14,
                                   equal_nan=True)


def test_resample_positions_are_uniform(oracle):
    m, draws = 37, []
    for b in range(60):
        draws.append(oracle.data_resample(3, b, 2, m * 20, seed=9) % m)
    counts = np.bincount(np.concatenate(draws), minlength=m)
    assert sps.chisquare(counts).pvalue > 1e-3
    # the four families and the genes draw different samples
    a = oracle.data_resample(0, 0, 2, 500, 1)
    assert not np.array_equal(a, oracle.data_resample(0, 0, 3, 500, 1))
    assert not np.array_equal(a, oracle.data_resample(1, 0, 2, 500, 1))
    assert not np.array_equal(a, oracle.data_resample(0, 1, 2, 500, 1))


def test_se_shrinks_with_cells(oracle):
    """The bootstrap SE of a mean over m cells scales like 1/sqrt(m)."""
    rng = np.random.default_rng(2)
    ses = []
    for n in (500, 8000):
        age = rng.integers(0, NC, n)
        cond = rng.integers(0, J, n)
        u = rng.poisson(4.0, (n, 1))
        l = rng.poisson(1.0, (n, 1))
        pulse = np.nonzero(cond < 6)[0]
        chase = np.nonzero(cond >= 6)[0]
        out = oracle.data_stats(u, l, age, cond, NC, J, pulse, chase, np.full((NC, J), 0.2),
                                n_boot=60)
        ses.append(out[0, 2 * NC:3 * NC].mean())           # pulse_mean_se
    ratio_se = ses[0] / ses[1]
    assert 2.5 < ratio_se < 6.0                             # sqrt(16) = 4
