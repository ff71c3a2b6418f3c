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


def test_multi_idx_call_equals_idx_calls(engine_factory):
    """abc_simulate_multi_idx_dev (subsets of each model's particles, one
    launch) against abc_simulate_idx_dev per model and against the same
    particles of a contiguous run: bit for bit."""
    import torch
    from abc_inference_transcription_amd import _lib
    eng = engine_factory(config_design(cells=200))
    L = eng.L
    dev = torch.device("cuda:0")
    s = torch.cuda.Stream(device=dev)
    eng.bind_stream(s)
    torch.cuda.set_stream(s)
    p = lambda t: C.c_void_p(t.data_ptr())
    K, KS = eng.n_stats, 2 * eng.nb * eng.J
    rng = np.random.default_rng(11)
    models, n_full, first = [2, 3, 5], 60, 1000
    full, pick, sep, mul = [], [], [], []
    for m in models:
        th = torch.empty(n_full * eng.n_params(m), dtype=torch.float64, device=dev)
        _lib.check(L.abc_sample_prior_dev(eng.ctx, m, 4, first, n_full, p(th)), eng.ctx, "prior")
        st = torch.empty(n_full * K, dtype=torch.float64, device=dev)
        sup = torch.empty(n_full * KS, dtype=torch.float64, device=dev)
        stp = torch.empty(n_full, dtype=torch.int64, device=dev)
        _lib.check(L.abc_simulate_dev(eng.ctx, m, 4, first, n_full, p(th), p(st), None, p(sup),
                                      p(stp)), eng.ctx, "simulate")
        full.append((th, sup, stp))
        keep = np.sort(rng.choice(n_full, size=int(rng.integers(5, 30)), replace=False))
        kd = torch.tensor(keep, dtype=torch.int64, device=dev)
        npar = eng.n_params(m)
        thk = th.view(npar, n_full)[:, kd].contiguous().view(-1)
        pidx = kd + first
        pick.append((keep, thk, pidx))
        mk = lambda n: dict(stats=torch.zeros(n * K, dtype=torch.float64, device=dev),
                            sup=torch.zeros(n * KS, dtype=torch.float64, device=dev),
                            steps=torch.zeros(n, dtype=torch.int64, device=dev))
        sep.append(mk(len(keep)))
        mul.append(mk(len(keep)))
    for m, (keep, thk, pidx), o in zip(models, pick, sep):
        _lib.check(L.abc_simulate_idx_dev(eng.ctx, m, 4, len(keep), p(pidx), p(thk),
                                          p(o["stats"]), None, p(o["sup"]), p(o["steps"])),
                   eng.ctx, "simulate_idx")
    VA = C.c_void_p * 3
    _lib.check(L.abc_simulate_multi_idx_dev(
        eng.ctx, 3, (C.c_int32 * 3)(*models), 4, (C.c_int64 * 3)(*[len(k) for k, _, _ in pick]),
        VA(*[q.data_ptr() for _, _, q in pick]), VA(*[t.data_ptr() for _, t, _ in pick]),
        VA(*[o["stats"].data_ptr() for o in mul]), None, VA(*[o["sup"].data_ptr() for o in mul]),
        VA(*[o["steps"].data_ptr() for o in mul])), eng.ctx, "simulate_multi_idx")
    torch.cuda.synchronize()
    for m, (keep, _, _), a, b, (th, sup, stp) in zip(models, pick, sep, mul, full):
        for k in ("stats", "sup", "steps"):
            x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
            assert np.array_equal(x.view(np.int64), y.view(np.int64)), (m, k)
        want = sup.view(KS, n_full)[:, torch.tensor(keep, device=dev)].cpu().numpy()
        got = b["sup"].view(KS, len(keep)).cpu().numpy()
        assert np.array_equal(np.isnan(want), np.isnan(got))
        ok = ~np.isnan(want)
        assert np.array_equal(want[ok].view(np.int64), got[ok].view(np.int64)), m
        assert np.array_equal(stp.cpu().numpy()[keep], b["steps"].cpu().numpy()), m
    with pytest.raises(RuntimeError):
        _lib.check(L.abc_simulate_multi_idx_dev(
            eng.ctx, 1, (C.c_int32 * 1)(2), 0, (C.c_int64 * 1)(4), None,
            (C.c_void_p * 1)(full[0][0].data_ptr()), (C.c_void_p * 1)(sep[0]["stats"].data_ptr()),
            None, None, None), eng.ctx, "simulate_multi_idx")
