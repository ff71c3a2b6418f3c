"""GPU parity through the C ABI: libabc.so against the CPU oracle on the same
seeded inputs.  Integer outputs (cell counts, age bins, work counters) and the
FP64 statistics / errors must be bit-identical; accepted sets identical."""
import numpy as np
import pytest

from abc_inference_transcription_amd.design import (CAPTURE_BETA, CAPTURE_EMPIRICAL,
                                                    config_design, reference_design)

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def empirical_design(golden_capture, cells=110):
    betas = golden_capture[:600]
    rng = np.random.default_rng(5)
    return reference_design(cells=cells, capture_mode=CAPTURE_EMPIRICAL, betas=betas,
                            beta_bin=rng.integers(0, 5, len(betas)),
                            beta_is_chase=rng.integers(0, 2, len(betas)))


def split_design(cells=120):
    d = config_design(cells=cells)
    d["split_target"], d["split_max"] = 200.0, 64  # force many sub-lanes per cell
    return d


DESIGNS = {
    "ref": lambda gc: reference_design(cells=110),
    "ref_empirical": lambda gc: empirical_design(gc),
    "config_beta": lambda gc: config_design(cells=180),
    "config_split": lambda gc: split_design(),
}


@pytest.mark.parametrize("dname", list(DESIGNS))
@pytest.mark.parametrize("model", [1, 2, 3, 4, 5])
def test_prior_bit_exact(oracle, engine_factory, golden_capture, dname, model):
    d = DESIGNS[dname](golden_capture)
    print(f"[prior] design={dname} model={model}", flush=True)
    eng = engine_factory(d)
    od = oracle.OracleDesign(d)
    got = eng.sample_prior(model, 257, seed=11, first=1000)
    want = oracle.sample_prior(od, model, 11, 1000, 257)
    assert same_bits(got, want)


@pytest.mark.parametrize("dname", list(DESIGNS))
@pytest.mark.parametrize("model", [1, 2, 3, 4, 5])
def test_simulate_bit_exact(oracle, engine_factory, golden_capture, dname, model):
    d = DESIGNS[dname](golden_capture)
    n = 10
    print(f"[simulate] design={dname} model={model} n={n} cells={d['cells']}", flush=True)
    eng = engine_factory(d)
    od = oracle.OracleDesign(d)
    theta = oracle.sample_prior(od, model, 3, 40, n)
    if dname == "config_beta":
        theta[:, -1] = np.clip(theta[:, -1], -1, 1)  # keep the oracle fast
    want = oracle.simulate(od, model, 3, 40, theta)
    got = eng.simulate(model, theta, seed=3, first=40, cells=True, moments=True, superset=True)
    assert np.array_equal(got["cell_u"], want["cell_u"])
    assert np.array_equal(got["cell_l"], want["cell_l"])
    assert np.array_equal(got["cell_bin"], want["cell_bin"])
    assert np.array_equal(got["steps"], want["steps"])
    assert same_bits(got["stats"], want["stats_ref"])
    assert same_bits(got["moments"], want["moments"])
    assert same_bits(got["superset"], want["superset"])


def test_simulate_split_invariance_gpu(engine_factory):
    d = reference_design(cells=110)
    eng = engine_factory(d)
    th = eng.sample_prior(4, 6, seed=2)
    a = eng.simulate(4, th, seed=2, first=0, cells=True)
    b = eng.simulate(4, th[3:], seed=2, first=3, cells=True)
    assert np.array_equal(a["cell_u"][3:], b["cell_u"])
    assert same_bits(a["stats"][3:], b["stats"])


def test_simulate_launch_shape_invariance(engine_factory):
    """Results do not depend on the persistent grid size (work stealing)."""
    d = config_design(cells=300)
    eng = engine_factory(d)
    th = eng.sample_prior(2, 8, seed=9)
    a = eng.simulate(2, th, seed=9, cells=True)
    eng.set_ssa_blocks(3)
    b = eng.simulate(2, th, seed=9, cells=True)
    eng.set_ssa_blocks(0)
    assert np.array_equal(a["cell_u"], b["cell_u"]) and np.array_equal(a["steps"], b["steps"])
    assert same_bits(a["stats"], b["stats"])


def _data_from(eng, stats, rng, G):
    pick = rng.integers(0, stats.shape[0], G)
    data = stats[pick] * rng.uniform(0.9, 1.1, (G, stats.shape[1]))
    se = np.abs(data) * rng.uniform(0.05, 0.3, data.shape)
    data[0, 3] = 0.0
    se[0, 3] = 0.0  # exercise the eps guard
    return data, se


def test_errors_bit_exact(oracle, engine_factory, golden_summary):
    d = reference_design(cells=110)
    eng = engine_factory(d)
    th = eng.sample_prior(2, 64, seed=1)
    sim = eng.simulate(2, th, seed=1)
    stats = sim["stats"]
    rng = np.random.default_rng(2)
    data, se = _data_from(eng, stats, rng, 37)
    want = oracle.errors(stats, data, se, eng.parts)
    got = eng.errors(stats, data, se)
    assert same_bits(got, want)
    # the reference's own data statistics (20 genes)
    g = golden_summary
    D = np.hstack([g[k] for k in ("pulse_mean", "pulse_ff", "chase_mean", "chase_ff", "ratio_data",
                                  "mean_corr_data", "corr_mean_data")])
    S = np.hstack([g[k] for k in ("pulse_mean_se", "pulse_ff_se", "chase_mean_se", "chase_ff_se",
                                  "ratio_se", "mean_corr_se", "corr_mean_se")])
    assert same_bits(eng.errors(stats, D, S), oracle.errors(stats, D, S, eng.parts))


@pytest.mark.parametrize("mode", [0, 1])
def test_accept_matches_fp64_oracle(oracle, engine_factory, mode):
    d = reference_design(cells=110)
    eng = engine_factory(d)
    th = eng.sample_prior(1, 300, seed=4)
    stats = eng.simulate(1, th, seed=4)["stats"]
    rng = np.random.default_rng(3)
    data, se = _data_from(eng, stats, rng, 23)
    err = oracle.errors(stats, data, se, eng.parts)
    eps, k = 2.0, 17
    got = eng.accept(stats, data, se, mode=mode, eps=eps, k=k, first_index=500)
    total = 0
    for g in range(data.shape[0]):
        if mode == 0:
            idx, e = oracle.accept_gene(err[:, g], eps)
        else:
            order = np.lexsort((np.arange(len(err)), err[:, g]))[:k]
            idx, e = order + 1, err[order, g]
        assert np.array_equal(got[g][0], idx + 500), g
        assert same_bits(got[g][1], e)
        total += len(idx)
    assert total > 0


def test_screen_bound_holds(oracle, engine_factory):
    """|approx - exact| <= bound for every pair, over stats and data spanning
    many decades (also validates the MFMA lane layout)."""
    d = config_design(cells=120)
    eng = engine_factory(d)
    rng = np.random.default_rng(7)
    n, G, K = 300, 150, 47
    parts = [5, 5, 5, 5, 9, 9, 9]
    scale = 10.0 ** rng.uniform(-3, 6, K)
    data = rng.uniform(0.2, 2.0, (G, K)) * scale
    se = np.abs(data) * rng.uniform(0.05, 0.3, (G, K))
    stats = data[rng.integers(0, G, n)] * rng.uniform(0.5, 1.5, (n, K))
    stats[:7] = data[:7]                       # exact hits: err = 0
    approx, bound = eng.screen_errors(stats, data, se, parts, clamp=1e300)
    exact = oracle.errors(stats, data, se, parts, clamp=1e300)
    gap = np.abs(approx.astype(np.float64) - exact)
    assert np.all(gap <= bound.astype(np.float64) * (1 + 1e-6) + 1e-6 * np.abs(exact)), \
        float(np.max(gap - bound))
    assert np.median(bound / np.maximum(exact, 1e-9)) < 0.1    # the bound is not vacuous


@pytest.mark.parametrize("mode", [0, 1])
def test_accept_stream_matches_fp64(oracle, engine_factory, mode):
    """Batches screened one after another give the FP64 sets of the whole run;
    the GEMM screen keeps the FP64 rescoring to a fraction of all pairs."""
    d = reference_design(cells=110)
    eng = engine_factory(d)
    th = eng.sample_prior(2, 900, seed=12)
    stats = eng.simulate(2, th, seed=12)["stats"]
    rng = np.random.default_rng(4)
    data, se = _data_from(eng, stats, rng, 31)
    err = oracle.errors(stats, data, se, eng.parts)
    eps, k = 1.5, 25
    batches = [(stats[a:a + 300], 7000 + a) for a in range(0, 900, 300)]
    got, ncand = eng.accept_stream(batches, data, se, mode=mode, eps=eps, k=k)
    for g in range(data.shape[0]):
        if mode == 0:
            idx, e = oracle.accept_gene(err[:, g], eps)
        else:
            order = np.lexsort((np.arange(len(err)), err[:, g]))[:k]
            idx, e = order + 1, err[order, g]
        assert np.array_equal(got[g][0], idx + 7000), g
        assert same_bits(got[g][1], e)
    if mode == 0:  # (top-k here sits at the clamp value 10, where every pair is a candidate)
        assert ncand < 0.5 * stats.shape[0] * data.shape[0]
