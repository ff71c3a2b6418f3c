"""Host logic of the surrogate-screened path (no GPU): the acceptance ceiling's
shares against brute force over every subset of small particle sets."""
import itertools

import numpy as np
import pytest

from abc_inference_transcription_amd.screened_abc import ceiling_shares


def _brute(s, w, R):
    """Fewest particles and least work of any subset keeping share R of the pairs."""
    n = len(s)
    best_n, best_w = n, float(np.sum(w))
    need = R * float(np.sum(s))
    for k in range(1, n + 1):
        for sub in itertools.combinations(range(n), k):
            idx = list(sub)
            if float(np.sum(s[idx])) >= need * (1 - 1e-12):
                best_n = min(best_n, k)
                best_w = min(best_w, float(np.sum(w[idx])))
    return best_n, best_w


@pytest.mark.parametrize("seed", range(12))
def test_ceiling_shares_against_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(4, 11))
    # many particles accept nothing; the rest carry a few expected pairs
    s = np.where(rng.random(n) < 0.4, 0.0, rng.gamma(0.7, 1.0, n))
    if s.sum() == 0:
        s[0] = 1.0
    w = rng.uniform(1.0, 50.0, n)
    out = ceiling_shares(s, w, recalls=(0.5, 0.9, 0.99))
    assert out["particles"] == n
    assert out["share_with_any_acceptance"] == pytest.approx((s > 0).mean())
    for R in (0.5, 0.9, 0.99):
        bn, bw = _brute(s, w, R)
        # fewest particles: exact
        assert out[f"share_needed_{R}"] == pytest.approx(bn / n), R
        # least work: the fractional bound never exceeds any subset's cost
        f = out[f"work_share_needed_{R}"]
        assert 0.0 < f <= bw / w.sum() + 1e-12, (R, f, bw / w.sum())
        assert out[f"work_gain_bound_{R}"] == pytest.approx(1.0 / f)


def test_ceiling_shares_fractional_bound_is_tight_when_items_divide():
    """Equal-ratio particles: the greedy prefix is optimal, and the bound
    equals the work share of the particles it keeps."""
    s = np.array([4.0, 2.0, 1.0, 1.0])
    w = np.array([8.0, 4.0, 2.0, 2.0])   # pairs per unit of work 0.5 for each
    out = ceiling_shares(s, w, recalls=(0.75,))
    # 0.75 of 8 pairs = 6: the first two particles, 12 of 16 units of work
    assert out["work_share_needed_0.75"] == pytest.approx(12 / 16)
    assert out["share_needed_0.75"] == pytest.approx(2 / 4)


def test_ceiling_shares_nothing_accepted():
    out = ceiling_shares(np.zeros(5), np.ones(5), recalls=(0.99,))
    assert out["share_needed_0.99"] == 1.0
    assert out["work_share_needed_0.99"] == pytest.approx(1.0)
