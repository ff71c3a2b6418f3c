"""Top-k acceptance of particles given by global index (abc_accept_add_idx_dev,
the path of the surrogate's survivors): per gene the k smallest (FP64 error,
global index) pairs among the rows added, bit for bit against the oracle's
error matrix, with the rows' global indices unsorted and the rows added over
two calls (scripts/accepted_particles.jl:12-32 with a top-k rule,
scripts/compute_errors.jl:45-70)."""
import ctypes as C

import numpy as np
import pytest

from abc_inference_transcription_amd.design import config_design

pytestmark = pytest.mark.gpu

K_TOP = 25


def expected_topk(err, gidx, k):
    """Per gene: 0-based global indices and errors of the k smallest (error,
    global index) pairs; NaN errors are never accepted."""
    out = []
    for g in range(err.shape[1]):
        col = err[:, g]
        ok = np.nonzero(~np.isnan(col))[0]
        order = np.lexsort((gidx[ok], col[ok]))[:k]
        sel = ok[order]
        out.append((gidx[sel], col[sel]))
    return out


def test_indexed_topk_matches_oracle(engine_factory, oracle):
    import torch
    from abc_inference_transcription_amd import _lib
    from abc_inference_transcription_amd import screened_abc as S
    eng = engine_factory(config_design(cells=1000))
    L = eng.L
    # genes: statistics of other particles, SE = |stat| U[0.05, 0.3]
    th = eng.sample_prior(2, 600, first=900_000)
    sup = eng.simulate(2, th, first=900_000, superset=True)["superset"]
    data = np.asfortranarray(sup[np.all(np.isfinite(sup), axis=1)][:60])
    assert data.shape[0] == 60
    se = np.asfortranarray(np.abs(data) * np.random.default_rng(1).uniform(0.05, 0.3, data.shape))
    G, KS = data.shape
    # candidate rows: 700 particles of model 3; 320 of them in a shuffled order
    th = eng.sample_prior(3, 700, first=40_000)
    rows = eng.simulate(3, th, first=40_000, superset=True)["superset"]
    rng = np.random.default_rng(5)
    pick = rng.choice(700, size=320, replace=False)          # unsorted
    sub = np.asfortranarray(rows[pick])
    gidx = (40_000 + pick).astype(np.int64)
    err = oracle.errors(sub, data, se, [45, 45])
    assert (~np.isnan(err)).sum(0).min() >= 4 * K_TOP   # every gene has a real choice
    want = expected_topk(err, gidx, K_TOP)

    parts = np.array([45, 45], np.int32)
    S._lib.check(L.abc_accept_begin(eng.ctx, KS, G, data.ctypes.data_as(C.c_void_p),
                                    se.ctypes.data_as(C.c_void_p),
                                    parts.ctypes.data_as(C.c_void_p), 2, 10.0, 1, 0.0, K_TOP,
                                    0.02), eng.ctx, "accept_begin")
    dev = torch.device("cuda:0")
    for a, b in ((0, 150), (150, 320)):                    # two calls: thresholds tighten
        blk = torch.tensor(np.asfortranarray(sub[a:b]).ravel(order="F"), dtype=torch.float64,
                           device=dev)
        p = np.ascontiguousarray(gidx[a:b])
        torch.cuda.synchronize()
        _lib.check(L.abc_accept_add_idx_dev(eng.ctx, b - a, C.c_void_p(blk.data_ptr()),
                                            p.ctypes.data_as(C.c_void_p)), eng.ctx, "add_idx")
        torch.cuda.synchronize()
    got = S._end(eng, G, G * K_TOP + 16)
    for g in range(G):
        gi, ge = got[g]
        wi, we = want[g]
        assert np.array_equal(gi, wi), g
        assert np.array_equal(np.ascontiguousarray(ge).view(np.uint64),
                              np.ascontiguousarray(we).view(np.uint64)), g
