"""Pin the reference-order FP64 errors (compute_errors.jl:30-70) and the
acceptance rule (accepted_particles.jl:20-22)."""
import math
from fractions import Fraction

import numpy as np

PARTS = [5, 5, 5, 5, 11, 11, 11]


def exact_error(d, se, s, parts):
    """Exact rational evaluation of the same expression (no rounding)."""
    K = len(d)
    tot = Fraction(0)
    off = 0
    s2 = Fraction(0.1) * Fraction(0.1)
    for p in parts:
        e = Fraction(0)
        for i in range(off, off + p):
            eps = Fraction(0) if se[i] + d[i] != 0 else Fraction(0.0001)
            dd, ss, sg = Fraction(d[i]), Fraction(s[i]), Fraction(se[i])
            e += (dd - ss) ** 2 / (sg * sg + s2 * dd * dd + eps)
        tot += e / K
        off += p
    return tot


def test_error_zero_when_equal(oracle, golden_summary):
    g = golden_summary
    d = np.concatenate([g["pulse_mean"][0], g["pulse_ff"][0], g["chase_mean"][0], g["chase_ff"][0],
                        g["ratio_data"][0], g["mean_corr_data"][0], g["corr_mean_data"][0]])
    se = np.concatenate([g["pulse_mean_se"][0], g["pulse_ff_se"][0], g["chase_mean_se"][0],
                         g["chase_ff_se"][0], g["ratio_se"][0], g["mean_corr_se"][0], g["corr_mean_se"][0]])
    assert oracle.error_pair(d, se, d, PARTS) == 0.0


def test_error_against_exact_arithmetic(oracle, golden_summary):
    """On the reference's real data statistics (20-gene slice) and perturbed
    model statistics, the FP64 value is within a few ulp of exact."""
    g = golden_summary
    rng = np.random.default_rng(0)
    for gi in range(6):
        d = np.concatenate([g["pulse_mean"][gi], g["pulse_ff"][gi], g["chase_mean"][gi], g["chase_ff"][gi],
                            g["ratio_data"][gi], g["mean_corr_data"][gi], g["corr_mean_data"][gi]])
        se = np.concatenate([g["pulse_mean_se"][gi], g["pulse_ff_se"][gi], g["chase_mean_se"][gi],
                             g["chase_ff_se"][gi], g["ratio_se"][gi], g["mean_corr_se"][gi], g["corr_mean_se"][gi]])
        s = d * rng.uniform(0.7, 1.3, 53)
        got = oracle.error_pair(d, se, s, PARTS, clamp=1e300)
        want = float(exact_error(d, se, s, PARTS))
        assert abs(got - want) <= 8 * math.ulp(want)


def test_error_epsilon_guard_and_clamp(oracle):
    """eps = 1e-4 only where se + d == 0; the result is clamped at 10."""
    d = np.zeros(53)
    se = np.zeros(53)
    s = np.zeros(53)
    s[0] = 0.01  # (0 - 0.01)^2 / 1e-4 / 53
    got = oracle.error_pair(d, se, s, PARTS, clamp=10.0)
    assert abs(got - (0.01 ** 2 / 1e-4) / 53) < 1e-15
    s[:] = 100.0
    assert oracle.error_pair(d, se, s, PARTS, clamp=10.0) == 10.0
    # one stat: closed form with sigma = 0.1
    got = oracle.error_pair(np.array([2.0]), np.array([0.5]), np.array([1.0]), [1], clamp=1e9)
    assert abs(got - 1.0 / (0.25 + 0.01 * 4.0)) < 1e-15


def test_accept_sorted_by_error_then_index(oracle):
    e = np.array([5.0, 1.0, 4.8, 1.0, 0.5, 4.80001, np.nan, 10.0])
    idx, err = oracle.accept_gene(e, 4.8)
    assert list(idx) == [5, 2, 4, 3]          # 1-based, ties by index, NaN never accepted
    assert list(err) == [0.5, 1.0, 1.0, 4.8]
    idx, _ = oracle.accept_gene(np.array([9.0, 7.0]), 4.8)
    assert len(idx) == 0


def test_error_matrix_layout(oracle):
    rng = np.random.default_rng(1)
    n, G, K = 7, 5, 53
    stats = rng.uniform(0, 2, (n, K))
    data = rng.uniform(0, 2, (G, K))
    se = rng.uniform(0.01, 0.3, (G, K))
    err = oracle.errors(stats, data, se, PARTS)
    for i in range(n):
        for g in range(G):
            assert err[i, g] == oracle.error_pair(data[g], se[g], stats[i], PARTS)
