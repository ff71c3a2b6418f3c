"""Surrogate-screened ABC on the GPU: the trained network keeps the accepted
pairs of the exhaustive plain path (recall) while skipping exact work."""
import pytest

pytestmark = pytest.mark.gpu


def test_screened_abc_recall_and_savings():
    import torch
    from abc_inference_transcription_amd import screened_abc as S
    from abc_inference_transcription_amd.design import config_design
    from abc_inference_transcription_amd.engine import Engine
    from abc_inference_transcription_amd.surrogate import Surrogate

    eng = Engine(config_design(cells=900))
    sur = Surrogate(90, max_batch=4096, seed=5)
    res = S.run(eng, sur, G=300, n_train=2000, n_eval=400, k=0, accept_rate=2e-4,
                z=[4.0, 3.0, 2.5], train_steps=800, batch=4096,
                log=lambda *a: print(*a, flush=True))
    print(res, flush=True)
    assert res["plain_pairs"] > 0
    # every screen setting is checked against the exhaustive path; the reported
    # one keeps recall >= 99 % and skips exact work the plain path spends
    assert res["recall"] >= 0.99
    assert res["ssa_work_fraction_avoided"] > 0.0
    assert all(r["recall"] <= 1.0 for r in res["by_z"])
    assert res["heldout_nll"] < res["heldout_nll_of_mean_predictor"]   # better than the prior mean
    assert 0.85 < res["heldout_coverage_2sigma"] < 1.0
    sur.close()
    eng.close()
