"""HIP-graph capture of the device-resident path and the SSA autotuner.

A captured prior -> simulate -> errors sequence must give, on every replay,
the bits of the same calls run eagerly; replays are interleaved with other
eager work and host synchronisations, and an in-place update to other
particles must match an eager run of those particles."""
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


@pytest.mark.parametrize("dname,model", [("config", 4), ("reference", 3)])
def test_graph_replays_match_eager(engine_factory, dname, model):
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
            for a, b in zip(step.outputs(), ref):
                np.testing.assert_array_equal(a, b)
        g = eng.capture(lambda: step(2000), graph=g)   # same sequence, other particles
        step.clear()
        eng.replay(g)
        torch.cuda.synchronize()
        for a, b in zip(step.outputs(), ref2):
            np.testing.assert_array_equal(a, b)
        eng.drop_graph(g)


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
    print(f"[autotune] blocks {blocks}, {ms:.2f} ms", flush=True)
    assert blocks > 0 and ms > 0
    assert eng.ssa_blocks() == blocks
    after = eng.simulate(5, th, first=77, superset=True)
    for k in ("stats", "superset", "steps"):
        np.testing.assert_array_equal(before[k], after[k])
