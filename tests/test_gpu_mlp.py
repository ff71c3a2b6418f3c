"""Surrogate kernels on the GPU: BF16-MFMA forward/backward against the float64
oracle, Adam training progress, and the screen's soundness."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def make(torch, ns=90, width=512, depth=4, max_batch=1024, seed=3):
    from abc_inference_transcription_amd.surrogate import Surrogate
    return Surrogate(ns, width=width, depth=depth, max_batch=max_batch, seed=seed)


def test_mlp_gradient_matches_fp64_oracle(torch):
    import mlp_ref as M
    s = make(torch)
    rng = np.random.default_rng(1)
    B = 512
    x = rng.normal(0, 0.7, (B, s.n_in)).astype(np.float32)
    t = rng.normal(0, 1, (B, 90)).astype(np.float32)
    xd = torch.tensor(x, device=s.dev)
    td = torch.tensor(t, device=s.dev)
    loss, g = s.grad(xd, td)
    theta = s.get_params().astype(np.float64)
    loss_ref, g_ref = M.loss_and_grad(theta, x.astype(np.float64), t.astype(np.float64),
                                      s.n_in, 512, 4, 90)
    assert abs(loss - loss_ref) < 2e-2 * abs(loss_ref)
    layers, _, _, _ = M.layout(s.n_in, 512, 4, 90)
    for (i, o, wo, bo) in layers:
        for a, b in ((g[wo:wo + i * o], g_ref[wo:wo + i * o]), (g[bo:bo + o], g_ref[bo:bo + o])):
            nrm = np.linalg.norm(b)
            if nrm == 0:
                assert np.linalg.norm(a) == 0
                continue
            rel = np.linalg.norm(a - b) / nrm
            cos = float(a @ b) / (np.linalg.norm(a) * nrm)
            assert rel < 6e-2 and cos > 0.998, (wo, rel, cos)
    s.close()


def test_mlp_count_outputs_gradient_matches_oracle(torch):
    """Poisson negative log-likelihood on the first statistics' log-rate
    outputs (the rest Gaussian): loss and gradients against the float64
    oracle."""
    import mlp_ref as M
    s = make(torch)
    nc = 3
    s.set_count_outputs(nc)
    rng = np.random.default_rng(7)
    B = 512
    x = rng.normal(0, 0.7, (B, s.n_in)).astype(np.float32)
    t = rng.normal(0, 1, (B, 90)).astype(np.float32)
    t[:, :nc] = rng.poisson(1.5, (B, nc))
    loss, g = s.grad(torch.tensor(x, device=s.dev), torch.tensor(t, device=s.dev))
    theta = s.get_params().astype(np.float64)
    loss_ref, g_ref = M.loss_and_grad(theta, x.astype(np.float64), t.astype(np.float64),
                                      s.n_in, 512, 4, 90, n_counts=nc)
    assert abs(loss - loss_ref) < 2e-2 * abs(loss_ref)
    layers, _, _, _ = M.layout(s.n_in, 512, 4, 90)
    for (i, o, wo, bo) in layers:
        for a, b in ((g[wo:wo + i * o], g_ref[wo:wo + i * o]), (g[bo:bo + o], g_ref[bo:bo + o])):
            nrm = np.linalg.norm(b)
            if nrm == 0:
                assert np.linalg.norm(a) == 0
                continue
            rel = np.linalg.norm(a - b) / nrm
            cos = float(a @ b) / (np.linalg.norm(a) * nrm)
            assert rel < 6e-2 and cos > 0.998, (wo, rel, cos)
    s.close()


def test_mlp_forward_matches_oracle(torch):
    import mlp_ref as M
    s = make(torch, max_batch=256)
    rng = np.random.default_rng(2)
    x = rng.normal(0, 0.7, (300, s.n_in)).astype(np.float32)   # two chunks of max_batch
    y = s.predict(torch.tensor(x, device=s.dev)).cpu().numpy()
    out, _, _ = M.forward(s.get_params().astype(np.float64), x.astype(np.float64), s.n_in, 512, 4, 90)
    ref = out[:, :180]
    assert np.max(np.abs(y - ref)) < 5e-2 * (np.max(np.abs(ref)) + 1)
    s.close()


def test_mlp_training_reduces_nll(torch):
    s = make(torch, max_batch=2048)
    rng = np.random.default_rng(4)
    n = 8192
    x = rng.uniform(-1, 1, (n, s.n_in)).astype(np.float32)
    A = rng.normal(0, 1, (s.n_in, 90)).astype(np.float32)
    mu = np.sin(x @ A / 3)
    t = (mu + 0.1 * rng.normal(0, 1, mu.shape)).astype(np.float32)
    xd, td = torch.tensor(x, device=s.dev), torch.tensor(t, device=s.dev)
    losses = s.train(xd, td, steps=300, batch=2048, lr=2e-3)
    assert losses[-20:].mean() < losses[:5].mean() - 1.0
    # the learned variance tracks the noise level 0.1 (log-variance ~ -4.6)
    y = s.predict(xd[:1024]).cpu().numpy()
    assert np.median(y[:, 90:]) < -2.5
    assert np.mean(np.abs(y[:, :90] - mu[:1024])) < 0.15
    s.close()


def test_gemm_tile_rows_change_speed_not_results(torch):
    """64- and 128-row and -column workgroup tiles accumulate every output in
    the same order: identical forward outputs and identical parameters after
    training."""
    rng = np.random.default_rng(9)
    n = 4096
    x = torch.tensor(rng.uniform(-1, 1, (n, 23)).astype(np.float32), device="cuda:0")
    t = torch.tensor(rng.normal(0, 1, (n, 90)).astype(np.float32), device="cuda:0")
    outs = []
    for rows, cols in ((128, 128), (64, 128), (64, 64), (128, 64), (0, 0)):
        s = make(torch, max_batch=2048)
        assert s.L.abc_mlp_set_tile_rows(s.h, rows) == 0
        assert s.L.abc_mlp_set_tile_cols(s.h, cols) == 0
        s.train(x, t, steps=12, batch=2048)
        outs.append((s.get_params(), s.predict(x).cpu().numpy()))
        s.close()
    for o in outs[1:]:
        assert np.array_equal(outs[0][0].view(np.int32), o[0].view(np.int32))
        assert np.array_equal(outs[0][1].view(np.int32), o[1].view(np.int32))
    s = make(torch)
    assert s.L.abc_mlp_set_tile_rows(s.h, 96) != 0
    s.close()


def test_training_is_reproducible_across_epochs_and_streams(torch):
    """Several epochs (each reshuffles into new torch buffers) give the same
    parameters on the network's own stream, bound to torch's null stream, and
    bound to a torch side stream: no step reads a buffer torch has reused."""
    rng = np.random.default_rng(12)
    n = 4096
    x = torch.tensor(rng.uniform(-1, 1, (n, 23)).astype(np.float32), device="cuda:0")
    t = torch.tensor(rng.normal(0, 1, (n, 90)).astype(np.float32), device="cuda:0")
    side = torch.cuda.Stream(device="cuda:0")
    params = []
    for mode in ("own", "null", "side"):
        s = make(torch, max_batch=1024)
        if mode == "null":
            s.bind_stream(torch.cuda.default_stream(torch.device("cuda:0")))
        if mode == "side":
            s.bind_stream(side)
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                s.train(x, t, steps=20, batch=1024)
        else:
            s.train(x, t, steps=20, batch=1024)
        torch.cuda.synchronize()
        params.append(s.get_params())
        s.close()
    assert np.all(np.isfinite(params[0]))
    for p in params[1:]:
        assert np.array_equal(params[0].view(np.int32), p.view(np.int32))
