"""The device's elementary functions against the oracle's, bit for bit
(abc_math_probe_dev): e^x in both device forms, over the normal and subnormal
result ranges and the special values, and the table-driven ln x."""
import ctypes as C

import numpy as np
import pytest

from abc_inference_transcription_amd.design import config_design

pytestmark = pytest.mark.gpu


def _probe(eng, kind, x):
    import torch
    from abc_inference_transcription_amd import _lib
    xd = torch.tensor(np.ascontiguousarray(x, np.float64), device="cuda:0")
    yd = torch.full_like(xd, 7.0)
    _lib.check(eng.L.abc_math_probe_dev(eng.ctx, kind, len(x), C.c_void_p(xd.data_ptr()),
                                        C.c_void_p(yd.data_ptr())), eng.ctx, "math_probe")
    torch.cuda.synchronize()
    return yd.cpu().numpy()


def _same_bits(got, want):
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    bad = np.flatnonzero(got[~nan].view(np.int64) != want[~nan].view(np.int64))
    return bad


def _exp_inputs(rng, top):
    return np.concatenate([
        rng.uniform(-760.0, top, 60_000),
        rng.uniform(-745.3, -700.0, 40_000),      # subnormal and smallest normal results
        rng.uniform(-1e-3, 1e-3, 10_000),
        -np.exp(rng.uniform(-700.0, 6.0, 10_000)),  # tiny to large negative arguments
        np.array([0.0, -0.0, -745.2, -745.19999, -708.39, -708.4, -1e-300, -np.inf]),
    ])


@pytest.mark.parametrize("kind", [0, 1])
def test_exp_matches_oracle(engine_factory, oracle, kind):
    eng = engine_factory(config_design(cells=100))
    rng = np.random.default_rng(40 + kind)
    x = _exp_inputs(rng, 709.78)
    if kind == 0:
        x = np.concatenate([x, [np.nan, np.inf, 709.79, 709.78, 800.0]])
    got = _probe(eng, kind, x)
    want = np.array([oracle.exp(v) for v in x])
    bad = _same_bits(got, want)
    assert len(bad) == 0, [(x[i], got[i], want[i]) for i in bad[:5]]
    sub = (want > 0) & (want < 2.2250738585072014e-308)
    assert sub.sum() > 10_000          # the subnormal range was exercised
    assert got[np.flatnonzero(x == -np.inf)][0] == 0.0


def test_ln_matches_oracle(engine_factory, oracle):
    eng = engine_factory(config_design(cells=100))
    rng = np.random.default_rng(7)
    x = np.concatenate([
        10.0 ** rng.uniform(-300.0, 300.0, 60_000),
        rng.uniform(0.5, 2.0, 20_000),
        1.0 + rng.uniform(-1e-6, 1e-6, 5_000),
        np.array([1.0, 2.0, 0.5, 1e-308, 1.7976931348623157e308]),
    ])
    got = _probe(eng, 2, x)
    want = np.array([oracle.log(v) for v in x])
    bad = _same_bits(got, want)
    assert len(bad) == 0, [(x[i], got[i], want[i]) for i in bad[:5]]


def test_probe_rejects_bad_arguments(engine_factory):
    from abc_inference_transcription_amd import _lib
    eng = engine_factory(config_design(cells=100))
    with pytest.raises(RuntimeError):
        _lib.check(eng.L.abc_math_probe_dev(eng.ctx, 3, 4, None, None), eng.ctx, "math_probe")
    with pytest.raises(RuntimeError):
        _lib.check(eng.L.abc_math_probe_dev(eng.ctx, 0, 4, None, None), eng.ctx, "math_probe")
    _lib.check(eng.L.abc_math_probe_dev(eng.ctx, 0, 0, None, None), eng.ctx, "math_probe")
