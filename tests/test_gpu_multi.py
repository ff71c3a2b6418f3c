"""abc_simulate_multi_dev (all models' particles in one burn-in and one SSA
launch) against one abc_simulate_dev call per model: statistics, superset,
moments and step counts bit for bit, with unequal and empty model batches."""
import ctypes as C

import numpy as np
import pytest

from abc_inference_transcription_amd.design import config_design

pytestmark = pytest.mark.gpu


def test_multi_model_call_equals_separate_calls(engine_factory):
    import torch
    from abc_inference_transcription_amd import _lib
    eng = engine_factory(config_design(cells=300))
    L = eng.L
    dev = torch.device("cuda:0")
    s = torch.cuda.Stream(device=dev)
    eng.bind_stream(s)
    torch.cuda.set_stream(s)
    models = [1, 2, 3, 4, 5]
    ns = [37, 0, 50, 23, 41]
    firsts = [5, 100, 0, 70000, 12]
    K, KS, KM = eng.n_stats, 2 * eng.nb * eng.J, 6 * eng.nb * eng.J
    p = lambda t: C.c_void_p(t.data_ptr())
    th, sep, mul = [], [], []
    for m, n, f in zip(models, ns, firsts):
        t = torch.empty(max(n, 1) * eng.n_params(m), dtype=torch.float64, device=dev)
        if n:
            _lib.check(L.abc_sample_prior_dev(eng.ctx, m, 3, f, n, p(t)), eng.ctx, "prior")
        th.append(t)
        mk = lambda: dict(stats=torch.zeros(max(n, 1) * K, dtype=torch.float64, device=dev),
                          sup=torch.zeros(max(n, 1) * KS, dtype=torch.float64, device=dev),
                          mom=torch.zeros(max(n, 1) * KM, dtype=torch.float64, device=dev),
                          steps=torch.zeros(max(n, 1), dtype=torch.int64, device=dev))
        sep.append(mk())
        mul.append(mk())
    for m, n, f, t, o in zip(models, ns, firsts, th, sep):
        _lib.check(L.abc_simulate_dev(eng.ctx, m, 3, f, n, p(t), p(o["stats"]), p(o["mom"]),
                                      p(o["sup"]), p(o["steps"])), eng.ctx, "simulate")
    VA = C.c_void_p * 5
    _lib.check(L.abc_simulate_multi_dev(
        eng.ctx, 5, (C.c_int32 * 5)(*models), 3, (C.c_int64 * 5)(*firsts), (C.c_int64 * 5)(*ns),
        VA(*[t.data_ptr() for t in th]), VA(*[o["stats"].data_ptr() for o in mul]),
        VA(*[o["mom"].data_ptr() for o in mul]), VA(*[o["sup"].data_ptr() for o in mul]),
        VA(*[o["steps"].data_ptr() for o in mul])), eng.ctx, "simulate_multi")
    torch.cuda.synchronize()
    for m, n, a, b in zip(models, ns, sep, mul):
        if n == 0:
            continue
        for k in ("stats", "sup", "mom", "steps"):
            x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
            assert np.array_equal(x.view(np.int64), y.view(np.int64)), (m, k)
    # argument errors
    with pytest.raises(RuntimeError):
        _lib.check(L.abc_simulate_multi_dev(eng.ctx, 17, None, 0, None, None, None, None, None,
                                            None, None), eng.ctx, "simulate_multi")
    with pytest.raises(RuntimeError):
        _lib.check(L.abc_simulate_multi_dev(
            eng.ctx, 1, (C.c_int32 * 1)(6), 0, (C.c_int64 * 1)(0), (C.c_int64 * 1)(4),
            (C.c_void_p * 1)(th[0].data_ptr()), (C.c_void_p * 1)(sep[0]["stats"].data_ptr()),
            None, None, None), eng.ctx, "simulate_multi")
