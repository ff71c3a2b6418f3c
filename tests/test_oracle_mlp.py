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


def test_mlp_oracle_count_outputs():
    """Count statistics (Poisson NLL on a log-rate output): the loss equals
    scipy's Poisson log-pmf up to the log t! constant inside [-30, 8] and its
    tangent at the bound outside, and the gradient matches central finite
    differences."""
    import mlp_ref as M
    from scipy.special import gammaln
    from scipy.stats import poisson
    rng = np.random.default_rng(5)
    n_in, width, depth, ns, nc = 23, 128, 2, 5, 2
    layers, n_par, in_pad, out_pad = M.layout(n_in, width, depth, ns)
    theta = np.zeros(n_par)
    for (i, o, wo, bo) in layers:
        theta[wo:wo + i * o] = rng.normal(0, 0.3, (o, i)).ravel()
        theta[bo:bo + o] = rng.normal(0, 0.1, o)
    B = 11
    x = rng.normal(0, 1, (B, n_in))
    t = rng.normal(0, 1, (B, ns))
    t[:, :nc] = rng.poisson(2.0, (B, nc))
    loss, g = M.loss_and_grad(theta, x, t, n_in, width, depth, ns, n_counts=nc)
    out, _, _ = M.forward(theta, x, n_in, width, depth, ns)
    # inside [-30, 8] the Poisson NLL itself; outside, its tangent at the bound
    o = out[:, :nc]
    ob = np.clip(o, -30.0, 8.0)
    assert np.any(o != ob) and np.any(o == ob)   # both branches are exercised
    pois = (-poisson.logpmf(t[:, :nc], np.exp(ob)) - gammaln(t[:, :nc] + 1)
            + (np.exp(ob) - t[:, :nc]) * (o - ob))
    lv = np.clip(out[:, ns + nc:2 * ns], -12, 12)
    gau = 0.5 * (lv + (t[:, nc:] - out[:, nc:ns]) ** 2 * np.exp(-lv))
    assert abs(loss - (pois.sum() + gau.sum()) / (B * ns)) < 1e-12 * max(1.0, abs(loss))
    # the log-variance slots of the count statistics carry no gradient
    i, o, wo, bo = layers[-1]
    Wg = g[wo:wo + i * o].reshape(o, i)
    assert np.all(Wg[ns:ns + nc] == 0) and np.all(g[bo + ns:bo + ns + nc] == 0)
    idx = []
    for (i, o, wo, bo) in layers:
        idx += [wo + rng.integers(0, min(i, n_in if wo == 0 else i)) + i * rng.integers(0, min(o, 2 * ns)) for _ in range(6)]
        idx += [bo + rng.integers(0, min(o, 2 * ns)) for _ in range(3)]
    h = 1e-4
    for p in idx:
        tp = theta.copy(); tp[p] += h
        tm = theta.copy(); tm[p] -= h
        fd = (M.loss_and_grad(tp, x, t, n_in, width, depth, ns, n_counts=nc)[0] -
              M.loss_and_grad(tm, x, t, n_in, width, depth, ns, n_counts=nc)[0]) / (2 * h)
        assert abs(fd - g[p]) <= 2e-6 * max(1.0, abs(fd)), (p, fd, g[p])
