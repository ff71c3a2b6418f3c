"""HIP-graph capture of the device-resident path and the SSA autotuner.

A captured prior -> simulate -> errors sequence must give, on every replay,
the bits of the same calls run eagerly and of the oracle's replay of the same
particles; replays are interleaved with other eager work and host
synchronisations, and an in-place update to other particles must match an
eager run of those particles.  A graph whose buffers a larger eager call
reallocated is refused on the host.  A captured top-k acceptance (one graph
updated per batch) gives the eager session's sets."""
import ctypes as C

import numpy as np
import pytest

from abc_inference_transcription_amd.design import config_design, reference_design

pytestmark = pytest.mark.gpu


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class Step:
    """Device buffers and the *_dev call sequence of one plain step."""

    def __init__(self, eng, torch, model, n, G, dev):
        self.eng, self.L, self.model, self.n, self.G = eng, eng.L, model, n, G
        np_, K = eng.n_params(model), eng.n_stats
        f64 = dict(dtype=torch.float64, device=dev)
        self.theta = torch.empty(n * np_, **f64)
        self.stats = torch.empty(n * K, **f64)
        self.sup = torch.empty(n * 2 * eng.nb * eng.J, **f64)
        self.steps = torch.empty(n, dtype=torch.int64, device=dev)
        self.err = torch.empty(n * G, **f64)
        rng = np.random.default_rng(3)
        data = rng.uniform(0.1, 5.0, (G, K))
        self.data = torch.tensor(data.ravel(order="F"), **f64)
        self.se = torch.tensor((data * 0.2).ravel(order="F"), **f64)
        self.parts = np.ascontiguousarray(eng.parts, np.int32)
        self.K = K

    def __call__(self, first):
        e, L = self.eng, self.L
        e._chk(L.abc_sample_prior_dev(e.ctx, self.model, 0, first, self.n, _ptr(self.theta)), "prior")
        e._chk(L.abc_simulate_dev(e.ctx, self.model, 0, first, self.n, _ptr(self.theta),
                                  _ptr(self.stats), None, _ptr(self.sup), _ptr(self.steps)), "sim")
        e._chk(L.abc_errors_dev(e.ctx, self.n, self.K, _ptr(self.stats), self.G, _ptr(self.data),
                                _ptr(self.se), self.parts.ctypes.data_as(C.c_void_p),
                                len(self.parts), 10.0, _ptr(self.err)), "errors")

    def outputs(self):
        return [t.cpu().numpy().copy() for t in (self.theta, self.stats, self.sup, self.steps, self.err)]

    def clear(self):
        for t in (self.theta, self.stats, self.sup, self.err):
            t.fill_(-1.0)
        self.steps.fill_(-1)


def same_bits(a, b):
    """Bit-identical where not NaN, NaN in the same places (IEEE leaves a NaN's
    payload to the hardware, and x86 and the GPU make different ones)."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint64), b[~nb].view(np.uint64))


def check_against_oracle(oracle, design, step, outs, first):
    """Replayed theta / stats / superset / steps / errors against the oracle's
    prior, simulation and FP64 errors of the same particles."""
    theta, stats, sup, steps, err = outs
    n, K, G = step.n, step.K, step.G
    od = oracle.OracleDesign(design)
    th = oracle.sample_prior(od, step.model, 0, first, n)
    assert same_bits(theta.reshape(-1, n).T, th)
    want = oracle.simulate(od, step.model, 0, first, th, cells=False)
    assert same_bits(stats.reshape(K, n).T, want["stats_ref"])
    assert same_bits(sup.reshape(-1, n).T, want["superset"])
    assert np.array_equal(steps, np.asarray(want["steps"], np.int64))
    data = step.data.cpu().numpy().reshape(K, G).T
    se = step.se.cpu().numpy().reshape(K, G).T
    e = oracle.errors(want["stats_ref"], data, se, step.parts)
    assert same_bits(err.reshape(G, n).T, e)


@pytest.mark.parametrize("dname,model", [("config", 4), ("reference", 3)])
def test_graph_replays_match_eager(engine_factory, oracle, dname, model):
    import torch
    dev = torch.device("cuda:0")
    design = config_design(cells=300) if dname == "config" else reference_design(cells=220)
    eng = engine_factory(design)
    s = torch.cuda.Stream(device=dev)
    eng.bind_stream(s)
    with torch.cuda.stream(s):
        step = Step(eng, torch, model, 45, 33, dev)
        other = Step(eng, torch, 2, 20, 33, dev)
        print(f"[graph] {dname} model {model}: eager", flush=True)
        step(1000)
        torch.cuda.synchronize()
        ref = step.outputs()
        step(2000)
        torch.cuda.synchronize()
        ref2 = step.outputs()
        g = eng.capture(lambda: step(1000))
        for rep in range(3):
            step.clear()
            other(5000 + rep)               # eager work between replays
            torch.cuda.synchronize()
            eng.replay(g)
            torch.cuda.synchronize()
            print(f"[graph] replay {rep}", flush=True)
            outs = step.outputs()
            for a, b in zip(outs, ref):
                np.testing.assert_array_equal(a, b)
            if rep == 0:
                check_against_oracle(oracle, design, step, outs, 1000)
        g = eng.capture(lambda: step(2000), graph=g)   # same sequence, other particles
        step.clear()
        eng.replay(g)
        torch.cuda.synchronize()
        outs = step.outputs()
        for a, b in zip(outs, ref2):
            np.testing.assert_array_equal(a, b)
        check_against_oracle(oracle, design, step, outs, 2000)
        eng.drop_graph(g)


def test_replay_refused_after_buffer_growth(engine_factory):
    """A larger eager call after the capture reallocates a buffer the graph
    holds: the replay is refused before any kernel runs, and a new capture
    works."""
    import torch
    dev = torch.device("cuda:0")
    eng = engine_factory(config_design(cells=150))
    s = torch.cuda.Stream(device=dev)
    eng.bind_stream(s)
    with torch.cuda.stream(s):
        small = Step(eng, torch, 2, 45, 7, dev)
        small(0)
        torch.cuda.synchronize()
        ref = small.outputs()
        g = eng.capture(lambda: small(0))
        big = Step(eng, torch, 2, 100, 7, dev)
        big(500)                                   # grows the context's buffers
        torch.cuda.synchronize()
        small.clear()
        with pytest.raises(RuntimeError, match="reallocated"):
            eng.replay(g)
        torch.cuda.synchronize()
        assert float(small.stats[0].item()) == -1.0   # nothing ran
        g = eng.capture(lambda: small(0), graph=g)
        eng.replay(g)
        torch.cuda.synchronize()
        for a, b in zip(small.outputs(), ref):
            np.testing.assert_array_equal(a, b)
        eng.drop_graph(g)


def test_captured_topk_acceptance_matches_eager(engine_factory):
    """Per batch, prior -> simulate -> top-k acceptance captured into one graph
    (updated in place for the next batch) and replayed; the session's sets
    equal those of the same batches run eagerly."""
    import torch
    dev = torch.device("cuda:0")
    eng = engine_factory(config_design(cells=600))
    s = torch.cuda.Stream(device=dev)
    eng.bind_stream(s)
    L, ctx = eng.L, eng.ctx
    KS = 2 * eng.nb * eng.J
    n, G, k, nbatch = 300, 40, 12, 4
    rng = np.random.default_rng(8)
    with torch.cuda.stream(s):
        th = torch.empty(n * eng.n_params(3), dtype=torch.float64, device=dev)
        st = torch.empty(n * eng.n_stats, dtype=torch.float64, device=dev)
        sup = torch.empty(n * KS, dtype=torch.float64, device=dev)
        # genes from the (finite) simulated statistics of other particles
        gt = eng.simulate(3, eng.sample_prior(3, 8 * G, first=90_000), first=90_000,
                          superset=True)["superset"]
        gt = gt[np.all(np.isfinite(gt), axis=1)][:G]
        assert len(gt) == G
        data = np.asfortranarray(gt * rng.uniform(0.9, 1.1, gt.shape))
        se = np.asfortranarray(np.abs(data) * rng.uniform(0.05, 0.3, data.shape))
        parts = np.array([KS // 2, KS // 2], np.int32)
        vp = lambda a: a.ctypes.data_as(C.c_void_p)

        def batch(b):
            eng._chk(L.abc_sample_prior_dev(ctx, 3, 0, b * n, n, _ptr(th)), "prior")
            eng._chk(L.abc_simulate_dev(ctx, 3, 0, b * n, n, _ptr(th), _ptr(st), None,
                                        _ptr(sup), None), "sim")
            eng._chk(L.abc_accept_add_dev(ctx, n, _ptr(sup), b * n), "add")

        def session(run):
            eng._chk(L.abc_accept_begin(ctx, KS, G, vp(data), vp(se), vp(parts), 2, 10.0, 1,
                                        0.0, k, 0.02), "begin")
            run()
            counts = np.zeros(G, np.int64)
            idx = np.zeros(G * k, np.int64)
            eo = np.zeros(G * k)
            eng._chk(L.abc_accept_end(ctx, vp(counts), vp(idx), vp(eo), G * k, None), "end")
            return counts, idx, eo

        def eager():
            for b in range(nbatch):
                batch(b)
        want = session(eager)
        graphs = {}

        def captured():
            g = None
            for b in range(nbatch):
                g = eng.capture(lambda: batch(b), graph=g)
                eng.replay(g)
            graphs["g"] = g
        got = session(captured)
        torch.cuda.synchronize()
        eng.drop_graph(graphs["g"])
    assert np.array_equal(want[0], got[0]) and int(want[0].sum()) > 0
    assert np.array_equal(want[1], got[1])
    assert same_bits(want[2], got[2])


def test_capture_refuses_host_calls_and_growth(engine_factory):
    import torch
    dev = torch.device("cuda:0")
    eng = engine_factory(config_design(cells=100))
    s = torch.cuda.Stream(device=dev)
    eng.bind_stream(s)
    with torch.cuda.stream(s):
        step = Step(eng, torch, 1, 8, 5, dev)
        step(0)
        torch.cuda.synchronize()
        big = Step(eng, torch, 1, 64, 5, dev)
        with pytest.raises(RuntimeError, match="graph_end|would grow|not allowed"):
            eng.capture(lambda: big(0))            # buffers would grow: refused
        with pytest.raises(RuntimeError):
            eng.capture(lambda: eng.sample_prior(1, 4))   # host-pointer call: refused
        # the context is usable afterwards, eagerly and captured
        step(0)
        torch.cuda.synchronize()
        ref = step.outputs()
        g = eng.capture(lambda: step(0))
        step.clear()
        eng.replay(g)
        torch.cuda.synchronize()
        for a, b in zip(step.outputs(), ref):
            np.testing.assert_array_equal(a, b)
        eng.drop_graph(g)


def test_autotune_changes_speed_not_results(engine_factory):
    eng = engine_factory(config_design(cells=400))
    th = eng.sample_prior(5, 40, first=77)
    before = eng.simulate(5, th, first=77, superset=True)
    blocks, ms = eng.autotune(models=(4, 5), n=64, first=10)
    print(f"[autotune] blocks {blocks}, claim {eng.ssa_claim()}, {ms:.2f} ms", flush=True)
    assert blocks > 0 and ms > 0
    assert eng.ssa_blocks() == blocks and eng.ssa_claim() in (1, 2, 4, 8)
    after = eng.simulate(5, th, first=77, superset=True)
    for k in ("stats", "superset", "steps"):
        np.testing.assert_array_equal(before[k], after[k])
    # every queue claim gives the same bits
    for claim in (2, 4, 8):
        eng.set_ssa_claim(claim)
        again = eng.simulate(5, th, first=77, superset=True)
        for k in ("stats", "superset", "steps"):
            np.testing.assert_array_equal(before[k], again[k])


def test_claim_walk_over_split_cells(engine_factory, oracle):
    """With a small split target most cells run as several sub-lanes, so a
    lane's claim steps across sub-lanes, cells and particles; every claim
    gives the bits of claim 1 and of the oracle."""
    design = dict(config_design(cells=120))
    design["split_target"] = 64.0
    eng = engine_factory(design)
    th = eng.sample_prior(4, 24, first=300)
    want = oracle.simulate(oracle.OracleDesign(design), 4, 0, 300, th, cells=False)
    for claim in (1, 2, 4, 8):
        eng.set_ssa_claim(claim)
        got = eng.simulate(4, th, first=300, superset=True)
        print(f"[claim] {claim}", flush=True)
        assert same_bits(got["superset"], want["superset"]), claim
        assert same_bits(got["stats"], want["stats_ref"]), claim
        assert np.array_equal(got["steps"], np.asarray(want["steps"], np.int64)), claim
