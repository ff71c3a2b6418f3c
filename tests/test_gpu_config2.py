"""The full-ABC-sweep workload's acceptance at its defining shape: per gene the
1,000 particles of smallest FP64 error (ties by index) out of 50,000 particles
of the five models (10,000 each; 2,500 cells, the configured cell design)
scored against 2,000 synthetic genes over the 90 superset statistics.  The
genes' data statistics are those of ground-truth particles drawn from the same
priors, with SE = |stat| * U[0.05, 0.3].  The streamed split-BF16 screen with
2 % margin, the FP64 rescoring and the on-device per-gene selection must give
each gene the oracle's list: the same indices in the same order, with
bit-identical errors (scripts/accepted_particles.jl:12-32 with a top-k rule,
scripts/compute_errors.jl:45-70)."""
import numpy as np
import pytest

from abc_inference_transcription_amd.design import config_design

pytestmark = pytest.mark.gpu

MODELS = (1, 2, 3, 4, 5)
PER_MODEL, G, K_TOP = 10_000, 2_000, 1_000


def oracle_topk(err, k):
    """Per gene: the k smallest (error, index) pairs of the FP64 matrix (NaN
    errors are never accepted); 1-based indices and the errors."""
    out = []
    for g in range(err.shape[1]):
        col = err[:, g]
        ok = np.nonzero(~np.isnan(col))[0]
        if len(ok) > k:
            v = np.partition(col[ok], k - 1)[k - 1]
            ok = ok[col[ok] <= v]
        order = np.lexsort((ok, col[ok]))[:k]
        sel = ok[order]
        out.append((sel + 1, col[sel]))
    return out


@pytest.fixture(scope="module")
def sweep():
    from abc_inference_transcription_amd.engine import Engine
    eng = Engine(config_design(cells=2500))
    rows = []
    for m in MODELS:
        th = eng.sample_prior(m, 420, first=3_000_000_000 + m * 1000)
        sup = eng.simulate(m, th, first=3_000_000_000 + m * 1000, superset=True)["superset"]
        rows.append(sup[np.all(np.isfinite(sup), axis=1)])
    data = np.vstack(rows)[:G]
    assert data.shape[0] == G
    se = np.abs(data) * np.random.default_rng(0).uniform(0.05, 0.3, data.shape)
    stats = []
    for m in MODELS:
        print(f"[config2] model {m}: {PER_MODEL} particles x 2500 cells", flush=True)
        th = eng.sample_prior(m, PER_MODEL, first=0)
        stats.append(eng.simulate(m, th, first=0, superset=True)["superset"])
    S = np.vstack(stats)
    yield eng, S, np.asfortranarray(data), np.asfortranarray(se)
    eng.close()


@pytest.fixture(scope="module")
def oracle_sets(oracle, sweep):
    eng, S, data, se = sweep
    print("[config2] oracle FP64 error matrix", flush=True)
    err = oracle.errors(S, data, se, [45, 45])
    return oracle_topk(err, K_TOP)


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def check(got, want):
    for g in range(G):
        gi, ge = got[g]
        wi, we = want[g]
        assert len(gi) == K_TOP and np.array_equal(gi, wi), g
        assert same_bits(ge, we), g


def test_topk_1000_streamed_per_model_blocks(sweep, oracle_sets):
    """Blocks of 4,096 rows per model (the benchmark's batches), indices of
    model m offset by (m - 1) * 10,000; the thresholds tighten block by block."""
    eng, S, data, se = sweep
    batches = []
    for q, m in enumerate(MODELS):
        rows = S[q * PER_MODEL:(q + 1) * PER_MODEL]
        for a in range(0, PER_MODEL, 4096):
            batches.append((rows[a:a + 4096], q * PER_MODEL + a))
    print("[config2] streamed top-1000 acceptance", flush=True)
    got, n_resc = eng.accept_stream(batches, data, se, parts=[45, 45], mode=1, k=K_TOP,
                                    margin_rel=0.02)
    check(got, oracle_sets)
    # each gene's threshold is the clamp (10) until its 1,000 best errors fall
    # below it, so early blocks are rescored whole; later ones only near it
    print(f"[config2] {n_resc} FP64 rescorings of {len(S) * G} pairs", flush=True)
    assert 0 < n_resc < len(S) * G


def test_topk_1000_one_batch(sweep, oracle_sets):
    """All 50,000 rows in one call: the session screens them in blocks of its
    own size (tens of thousands of rows), the first with no threshold yet."""
    eng, S, data, se = sweep
    got = eng.accept(S, data, se, parts=[45, 45], mode=1, k=K_TOP)
    check(got, oracle_sets)
