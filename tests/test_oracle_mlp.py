"""Pin the float64 surrogate oracle by central finite differences."""
import numpy as np


def test_mlp_oracle_gradient_matches_finite_differences():
    import mlp_ref as M
    rng = np.random.default_rng(0)
    n_in, width, depth, ns = 23, 128, 2, 7
    layers, n_par, in_pad, out_pad = M.layout(n_in, width, depth, ns)
    theta = np.zeros(n_par)
    for (i, o, wo, bo) in layers:
        W = rng.normal(0, 0.3, (o, i))
        theta[wo:wo + i * o] = W.ravel()
        theta[bo:bo + o] = rng.normal(0, 0.1, o)
    x = rng.normal(0, 1, (9, n_in))
    t = rng.normal(0, 1, (9, ns))
    loss, g = M.loss_and_grad(theta, x, t, n_in, width, depth, ns)
    # probe coordinates in every layer (weights and biases)
    idx = []
    for (i, o, wo, bo) in layers:
        idx += [wo + rng.integers(0, min(i, n_in if wo == 0 else i)) + i * rng.integers(0, min(o, 2 * ns)) for _ in range(6)]
        idx += [bo + rng.integers(0, min(o, 2 * ns)) for _ in range(3)]
    h = 1e-4  # truncation ~h^2, cancellation ~eps*|loss|/h: both below 1e-7 here
    for p in idx:
        tp = theta.copy(); tp[p] += h
        tm = theta.copy(); tm[p] -= h
        fd = (M.loss_and_grad(tp, x, t, n_in, width, depth, ns)[0] -
              M.loss_and_grad(tm, x, t, n_in, width, depth, ns)[0]) / (2 * h)
        assert abs(fd - g[p]) <= 2e-6 * max(1.0, abs(fd)), (p, fd, g[p])
    assert np.isfinite(loss)
