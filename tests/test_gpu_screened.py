"""Surrogate-screened ABC on the GPU: the trained network keeps the accepted
pairs of the exhaustive plain path (recall) while skipping exact work, and the
gain is measured against the plain path on the same held-out particles."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def screened():
    from abc_inference_transcription_amd import screened_abc as S
    from abc_inference_transcription_amd.design import config_design
    from abc_inference_transcription_amd.engine import Engine
    from abc_inference_transcription_amd.surrogate import Surrogate

    # the benchmark's shape, smaller: 2,000 genes accepting 4e-5 of the pairs,
    # 60,000 training particles, 10,000 calibration and 10,000 held-out ones
    eng = Engine(config_design(cells=2500))
    sur = Surrogate(90, max_batch=8192, seed=5)
    res = S.run(eng, sur, G=2000, n_train=12000, n_eval=2000, k=0, accept_rate=4e-5,
                z=[3.0, 2.5], train_steps=2500, batch=8192, n_cal=2000, cal_recall=0.995,
                direct_steps=1500, log=lambda *a: print(*a, flush=True))
    print({k: v for k, v in res.items() if k not in ("by_z", "by_rule")}, flush=True)
    yield eng, sur, res
    sur.close()
    eng.close()


def test_screened_abc_recall_and_savings(screened):
    _, _, res = screened
    assert res["plain_pairs"] > 0
    # every screen setting is checked against the exhaustive path
    assert all(r["recall"] <= 1.0 for r in res["by_z"])
    assert res["heldout_nll"] < res["heldout_nll_of_mean_predictor"]   # better than the prior mean
    assert 0.85 < res["heldout_coverage_2sigma"] < 1.0
    # rules calibrated on other particles keep the held-out recall and skip
    # exact work: what the trained network gains over simulating every particle
    rules = res["by_rule"]
    assert {r["rule"] for r in rules} == {"expected_genes", "max_margin",
                                          "direct_expected_genes", "direct_genes_per_step",
                                          "blended_genes_per_step"}
    best = max(rules, key=lambda r: r["ssa_work_fraction_avoided"] if r["recall"] >= 0.98 else -1)
    assert best["recall"] >= 0.98, rules
    assert best["ssa_work_fraction_avoided"] > 0.2, rules
    # measured about 3.4x at this size (0.62 s plain, 0.18 s screened)
    assert best["speedup_accepted_per_s"] > 1.8, rules


def test_score_rule_reproduces_pass_flags(screened):
    """abc_screen_dev passes a particle when its largest standardised margin
    is >= -z; abc_screen_score_dev reports that margin."""
    import ctypes as C
    from abc_inference_transcription_amd import screened_abc as S
    eng, sur, _ = screened
    L = eng.L
    th, sup, _ = S.simulate_block(eng, 3, 500, 123_000)
    rows = sup.view(90, 500).t().cpu().numpy()
    data = np.asfortranarray(rows[np.all(np.isfinite(rows), axis=1)][:40] * 1.1)
    assert data.shape[0] == 40
    se = np.asfortranarray(np.abs(data) * 0.2 + 1e-3)
    parts = np.array([45, 45], np.int32)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    eng._chk(L.abc_accept_begin(eng.ctx, 90, 40, vp(data), vp(se), vp(parts), 2, 10.0, 0, 0.5,
                                0, 0.0), "begin")
    x = sur.features(3, th)
    sc = sur.scores(eng, x).cpu().numpy()
    for z in (2.0, 3.0):
        pas, _ = sur.screen(eng, x, z=z)
        flags = pas.cpu().numpy().astype(bool)
        far = np.abs(sc[:, 1] + z) > 1e-3
        assert far.mean() > 0.9
        assert np.array_equal(flags[far], (sc[:, 1] >= -z)[far]), z
    assert np.all(sc[:, 0] >= 0) and np.all(sc[:, 0] <= 40 + 1e-3)
    S._end(eng, 40, 16)
