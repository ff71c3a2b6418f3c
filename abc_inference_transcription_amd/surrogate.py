"""Surrogate-screened ABC: train the MLP on simulator output and let it pick
which prior particles get the exact simulation.

All compute runs in libabc.so (features, targets, forward/backward, Adam,
screen); torch only supplies device buffers and the stream.  No reference
counterpart: the reference simulates every prior particle
(scripts/abc_simulation.jl:88-97).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class Surrogate:
    """(model, log10 parameters) -> mean and log-variance of every statistic
    in the standardised asinh space."""

    def __init__(self, n_stats: int, n_bins: int = 5, width: int = 512, depth: int = 4,
                 max_batch: int = 8192, seed: int = 0, device: int = 0, stream=None):
        import torch
        self.torch = torch
        self.L = _lib.load()
        self.n_stats, self.n_bins = n_stats, n_bins
        self.n_in = 5 + 3 * n_bins + 3
        self.max_batch = max_batch
        self.dev = torch.device(f"cuda:{device}")
        h = C.c_void_p()
        rc = self.L.abc_mlp_create(C.byref(h), device, self.n_in, width, depth, n_stats,
                                   max_batch, seed)
        if rc != 0:
            raise RuntimeError(f"abc_mlp_create failed ({rc})")
        self.h = h
        self.stream = None
        if stream is not None:
            self.bind_stream(stream)
        self.mean = torch.zeros(n_stats, dtype=torch.float32, device=self.dev)
        self.sd = torch.ones(n_stats, dtype=torch.float32, device=self.dev)

    def bind_stream(self, torch_stream):
        """Launch every later call on this torch stream.  The null stream
        cannot be shared (the library then keeps a stream of its own), so
        binding it leaves the network unbound and every call orders itself
        against torch's stream by synchronising."""
        self._chk(self.L.abc_mlp_set_stream(self.h, C.c_void_p(torch_stream.cuda_stream)),
                  "set_stream")
        self.stream = torch_stream if torch_stream.cuda_stream != 0 else None

    def _chk(self, rc, what):
        if rc != 0:
            m = self.L.abc_mlp_last_error(self.h)
            raise RuntimeError(f"{what} failed ({rc}): {m.decode() if m else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.L.abc_mlp_destroy(self.h)
            self.h = None

    def set_count_outputs(self, n_counts: int):
        """The first n_counts statistics become counts (Poisson NLL on a
        log-rate output: predict()'s mean column holds log E[count])."""
        self._chk(self.L.abc_mlp_set_count_outputs(self.h, n_counts), "set_count_outputs")

    @property
    def n_params(self) -> int:
        return self.L.abc_mlp_n_params(self.h)

    def _s(self):
        return C.c_void_p(self.stream.cuda_stream) if self.stream is not None else None

    def _after_torch(self):
        """Order the next network call after the torch work that made its
        inputs.  Unless bound to torch's current stream, the network runs on
        a stream of its own that does not wait for torch's."""
        cur = self.torch.cuda.current_stream(self.dev)
        if self.stream is None or self.stream.cuda_stream != cur.cuda_stream:
            cur.synchronize()

    def _before_torch(self):
        """Wait for the network's queued work before torch reuses memory it
        reads (only needed when the network has a stream of its own)."""
        cur = self.torch.cuda.current_stream(self.dev)
        if self.stream is None or self.stream.cuda_stream != cur.cuda_stream:
            self._chk(self.L.abc_mlp_synchronize(self.h), "sync")

    # ---------------------------------------------------------------- data
    def features(self, model: int, theta_dev):
        """theta_dev: torch float64 holding n x n_params column-major."""
        n = theta_dev.numel() // (6 if model in (1, 2) else 10)
        x = self.torch.empty((n, self.n_in), dtype=self.torch.float32, device=self.dev)
        self._after_torch()
        rc = self.L.abc_features_dev(model, n, _ptr(theta_dev), self.n_bins, _ptr(x), self._s())
        self._chk(rc, "features")
        return x

    def fit_normalizer(self, stats_np: np.ndarray):
        """Per-statistic mean / sd of asinh(stat) over a training set (n x K)."""
        a = np.arcsinh(np.where(np.isfinite(stats_np), stats_np, 0.0))
        self.mean = self.torch.tensor(a.mean(0), dtype=self.torch.float32, device=self.dev)
        self.sd = self.torch.tensor(np.maximum(a.std(0), 1e-3), dtype=self.torch.float32,
                                    device=self.dev)

    def targets(self, stats_dev, n: int):
        t = self.torch.empty((n, self.n_stats), dtype=self.torch.float32, device=self.dev)
        self._after_torch()
        rc = self.L.abc_targets_dev(n, _ptr(stats_dev), self.n_stats, _ptr(self.mean),
                                    _ptr(self.sd), _ptr(t), self._s())
        self._chk(rc, "targets")
        return t

    # ------------------------------------------------------------ training
    def train(self, x, t, steps: int, batch: int = 8192, lr: float = 1e-3, beta1: float = 0.9,
              beta2: float = 0.999, eps: float = 1e-8, weight_decay: float = 0.0,
              log_every: int = 0, seed: int = 0):
        torch = self.torch
        n = x.shape[0]
        batch = min(batch, self.max_batch, n)
        g = torch.Generator(device=self.dev)
        g.manual_seed(seed)
        loss_buf = torch.zeros(steps, dtype=torch.float32, device=self.dev)
        perm, pos = None, n
        for s in range(steps):
            if pos + batch > n:
                # the previous epoch's shuffled copies may still be read by
                # steps queued on the network's stream: let them finish before
                # torch may reuse that memory
                self._before_torch()
                perm = torch.randperm(n, device=self.dev, generator=g)
                xs, ts = x[perm].contiguous(), t[perm].contiguous()
                self._after_torch()
                pos = 0
            xb, tb = xs[pos:pos + batch], ts[pos:pos + batch]
            pos += batch
            lr_s = lr * 0.5 * (1.0 + np.cos(np.pi * s / steps))  # cosine schedule
            rc = self.L.abc_mlp_train_step_dev(self.h, batch, _ptr(xb), _ptr(tb), lr_s, beta1,
                                               beta2, eps, weight_decay, _ptr(loss_buf[s:s + 1]))
            self._chk(rc, "train_step")
            if log_every and (s + 1) % log_every == 0:
                print(f"[surrogate] step {s + 1}/{steps} loss {float(loss_buf[s].item()):.4f}",
                      flush=True)
        self._chk(self.L.abc_mlp_synchronize(self.h), "sync")
        return loss_buf.cpu().numpy()

    def predict(self, x):
        y = self.torch.empty((x.shape[0], 2 * self.n_stats), dtype=self.torch.float32,
                             device=self.dev)
        self._after_torch()
        self._chk(self.L.abc_mlp_forward_dev(self.h, x.shape[0], _ptr(x), _ptr(y)), "forward")
        self._chk(self.L.abc_mlp_synchronize(self.h), "sync")
        return y

    def grad(self, x, t):
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        self._after_torch()
        rc = self.L.abc_mlp_grad_dev(self.h, x.shape[0], _ptr(x), _ptr(t),
                                     g.ctypes.data_as(C.c_void_p), C.byref(loss))
        self._chk(rc, "grad")
        return loss.value, g

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        self._chk(self.L.abc_mlp_get_params(self.h, p.ctypes.data_as(C.c_void_p)), "get_params")
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        self._chk(self.L.abc_mlp_set_params(self.h, p.ctypes.data_as(C.c_void_p)), "set_params")

    def scores(self, eng, x, margin: float = 0.0):
        """Per particle (torch float32, n x 2) against the engine's open
        acceptance session: expected accepting genes, largest standardised
        margin."""
        n = x.shape[0]
        sc = self.torch.empty((n, 2), dtype=self.torch.float32, device=self.dev)
        self._after_torch()
        rc = self.L.abc_screen_score_dev(eng.ctx, self.h, n, _ptr(x), _ptr(self.mean),
                                         _ptr(self.sd), margin, _ptr(sc))
        _lib.check(rc, eng.ctx, "screen_score")
        return sc

    def screen(self, eng, x, z: float = 4.0, margin: float = 0.0):
        """pass flags (torch uint8) against the engine's open acceptance session."""
        n = x.shape[0]
        pas = self.torch.zeros(n, dtype=self.torch.uint8, device=self.dev)
        npass = C.c_int64()
        self._after_torch()
        rc = self.L.abc_screen_dev(eng.ctx, self.h, n, _ptr(x), _ptr(self.mean), _ptr(self.sd),
                                   z, margin, _ptr(pas), C.byref(npass))
        _lib.check(rc, eng.ctx, "screen")
        return pas, npass.value
