"""The data-side summary statistics on the GPU (abc_data_stats: get_summary_stats
of scripts/data_summary_statistics.jl for every gene, with bootstrap SEs)
against the CPU restatement: bit for bit, since the cell sums are exact
integers and every floating-point step runs in the same order."""
import ctypes as C

import numpy as np
import pytest

from test_oracle_data import NC, J, edge_case, synthetic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng(engine_factory):
    from abc_inference_transcription_amd.design import reference_design
    return engine_factory(reference_design(cells=55))


def _same_bits(a, b):
    """equal bits, except that a NaN matches a NaN of either sign (0/0 gives
    a positive NaN on the GPU and a negative one on x86)"""
    a, b = np.ascontiguousarray(a, np.float64), np.ascontiguousarray(b, np.float64)
    if a.shape != b.shape or not np.array_equal(np.isnan(a), np.isnan(b)):
        return False
    m = ~np.isnan(a)
    return np.array_equal(a[m].view(np.int64), b[m].view(np.int64))


@pytest.mark.parametrize("case,n_boot,seed", [("synthetic", 20, 0), ("synthetic", 100, 3),
                                               ("edge", 50, 1)])
def test_data_stats_bit_exact(oracle, eng, case, n_boot, seed):
    u, l, age, cond, pulse, chase, w = synthetic(seed, n=600, G=9) if case == "synthetic" else edge_case()
    print(f"\n[data_stats] {case} n={len(age)} G={u.shape[1]} B={n_boot}", flush=True)
    got = eng.data_stats(u, l, age, cond, NC, J, pulse, chase, w, n_boot=n_boot, seed=seed)
    want = oracle.data_stats(u, l, age, cond, NC, J, pulse, chase, w, n_boot=n_boot, seed=seed)
    assert _same_bits(got, want), np.nanmax(np.abs(got - want))


def test_data_stats_reference_shape(oracle, eng):
    """The reference's shape per gene (5 age clusters, 11 conditions, 100
    resamples) over 5,422 cells with counts of the size of the data, for 24
    genes; ragged pulse and chase sets (not every cell in a condition)."""
    rng = np.random.default_rng(12)
    n, G = 5422, 24
    age = rng.integers(0, NC, n)
    cond = rng.integers(-1, J, n)
    u = rng.negative_binomial(2, rng.uniform(0.05, 0.6, G), (n, G))
    l = rng.negative_binomial(1, rng.uniform(0.2, 0.9, G), (n, G))
    pulse = np.nonzero((cond >= 0) & (cond <= 5))[0]
    chase = np.nonzero(cond >= 6)[0]
    w = rng.random((NC, J))
    w /= w.sum()
    print(f"\n[data_stats] reference shape n={n} G={G}", flush=True)
    got = eng.data_stats(u, l, age, cond, NC, J, pulse, chase, w, n_boot=100, seed=77)
    want = oracle.data_stats(u, l, age, cond, NC, J, pulse, chase, w, n_boot=100, seed=77)
    assert _same_bits(got, want)
    assert np.all(np.isfinite(got))


def test_data_stats_many_launches(oracle, eng):
    """A resample count whose scratch exceeds one launch's (1 GiB) splits the
    genes over several launches; the result does not depend on the split."""
    rng = np.random.default_rng(5)
    n, G, nc, nj = 9, 13, 8, 16
    age = rng.integers(0, nc, n)
    cond = rng.integers(0, nj, n)
    u = rng.poisson(2.0, (n, G))
    l = rng.poisson(1.0, (n, G))
    pulse = np.arange(0, n, 2)
    chase = np.arange(1, n, 2)
    w = np.full((nc, nj), 1.0 / nc)
    print("\n[data_stats] several launches (n_boot 2^17: 12 genes per launch)", flush=True)
    got = eng.data_stats(u, l, age, cond, nc, nj, pulse, chase, w, n_boot=1 << 17, seed=2)
    want = oracle.data_stats(u, l, age, cond, nc, nj, pulse, chase, w, n_boot=1 << 17, seed=2)
    assert _same_bits(got, want)


def test_data_stats_dev_rejects_large_counts(eng):
    """abc_data_stats_dev: counts in device memory are checked by the kernel."""
    import torch
    u, l, age, cond, pulse, chase, w = synthetic(1, n=50, G=2)
    dev = torch.device("cuda:0")
    ut = torch.tensor(np.asfortranarray(u, np.uint32).T.reshape(-1).astype(np.int64), device=dev).to(torch.int32)
    lt = torch.tensor(np.asfortranarray(l, np.uint32).T.reshape(-1).astype(np.int64), device=dev).to(torch.int32)
    out = torch.zeros(2 * (8 * NC + 6 * J), dtype=torch.float64, device=dev)
    a = np.ascontiguousarray(age, np.int32); c = np.ascontiguousarray(cond, np.int32)
    p = np.ascontiguousarray(pulse, np.int32); ch = np.ascontiguousarray(chase, np.int32)
    wf = np.asfortranarray(w, np.float64)
    vp = lambda x: C.c_void_p(x.ctypes.data)

    def call():
        return eng.L.abc_data_stats_dev(eng.ctx, 2, 50, C.c_void_p(ut.data_ptr()),
                                        C.c_void_p(lt.data_ptr()), NC, vp(a), J, vp(c), vp(p),
                                        len(p), vp(ch), len(ch), vp(wf), 10, 0,
                                        C.c_void_p(out.data_ptr()))
    assert call() == 0
    host = np.ascontiguousarray(out.cpu().numpy().reshape(8 * NC + 6 * J, 2).T)
    want = eng.data_stats(u, l, age, cond, NC, J, pulse, chase, w, n_boot=10, seed=0)
    assert _same_bits(host, want)
    ut[7] = 1 << 20
    assert call() != 0


def test_data_stats_rejects_bad_arguments(eng):
    u, l, age, cond, pulse, chase, w = synthetic(2, n=40, G=1)
    ok = dict(n_boot=5)
    with pytest.raises(RuntimeError):
        eng.data_stats(u, l, np.where(age == 0, NC, age), cond, NC, J, pulse, chase, w, **ok)
    with pytest.raises(RuntimeError):
        eng.data_stats(u, l, age, np.where(cond == 0, J, cond), NC, J, pulse, chase, w, **ok)
    with pytest.raises(RuntimeError):
        eng.data_stats(u, l, age, cond, NC, J, np.append(pulse, 40), chase, w, **ok)
    with pytest.raises(RuntimeError):
        eng.data_stats(u + (1 << 20), l, age, cond, NC, J, pulse, chase, w, **ok)
    with pytest.raises(RuntimeError):
        eng.data_stats(u, l, age, cond, NC, J, pulse, chase, w, n_boot=1)
    with pytest.raises(RuntimeError):
        eng.data_stats(u, l, age % 9, cond, 9, J, pulse, chase, np.ones((9, J)), **ok)
    w2 = w.copy(); w2[0, 0] = np.nan
    with pytest.raises(RuntimeError):
        eng.data_stats(u, l, age, cond, NC, J, pulse, chase, w2, **ok)
    # no genes, and no cells
    assert eng.data_stats(np.zeros((40, 0)), np.zeros((40, 0)), age, cond, NC, J, pulse, chase, w,
                          **ok).shape == (0, 8 * NC + 6 * J)
    empty = eng.data_stats(np.zeros((0, 2)), np.zeros((0, 2)), [], [], NC, J, [], [], w, **ok)
    assert empty.shape == (2, 8 * NC + 6 * J) and np.all(empty == 0.0)


def test_data_stats_large_family_path(oracle, eng):
    """Families of more than 16,384 cells take the per-draw path
    (data_stats_kernel); smaller ones count draws per cell first
    (data_stats_sorted_kernel), up to 16,384 cells per family.  Both equal
    the CPU restatement."""
    for n, B in ((16500, 6), (12000, 10)):
        u, l, age, cond, pulse, chase, w = synthetic(8, n=n, G=2)
        print(f"\n[data_stats] {n} cells", flush=True)
        got = eng.data_stats(u, l, age, cond, NC, J, pulse, chase, w, n_boot=B, seed=4)
        want = oracle.data_stats(u, l, age, cond, NC, J, pulse, chase, w, n_boot=B, seed=4)
        assert _same_bits(got, want), n


def test_data_stats_whole_lanes(oracle, eng):
    """Families whose sizes are multiples of 256 (every lane of the packed
    layout holds cells, no padding) and one of fewer than 256 cells (at most
    one cell per lane)."""
    rng = np.random.default_rng(21)
    n, G = 768, 4
    age = rng.integers(0, NC, n)
    cond = rng.integers(0, J, n)
    u = rng.poisson(3.0, (n, G))
    l = rng.poisson(1.0, (n, G))
    pulse = np.sort(rng.choice(n, 512, replace=False))
    chase = np.sort(rng.choice(n, 200, replace=False))
    w = rng.random((NC, J))
    w /= w.sum(axis=0)
    print(f"\n[data_stats] whole lanes n={n} pulse=512 chase=200", flush=True)
    got = eng.data_stats(u, l, age, cond, NC, J, pulse, chase, w, n_boot=17, seed=6)
    want = oracle.data_stats(u, l, age, cond, NC, J, pulse, chase, w, n_boot=17, seed=6)
    assert _same_bits(got, want)
