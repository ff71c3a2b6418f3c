"""Float64 restatement of the surrogate network (TEST INFRASTRUCTURE ONLY).

Same architecture and parameter layout as libabc's abc_mlp: per layer
W (out_pad x in_pad, row-major) then b (out_pad); hidden layers SiLU; the
output row is [mean (ns) | log-variance (ns) | padding]; the loss is the mean
Gaussian negative log-likelihood over the batch and the ns statistics,
0.5 * (lv + (t - mu)^2 exp(-lv)) with lv clipped to [-12, 12].  The first
n_counts statistics may instead be counts: their mean slot holds a log-rate
o with the Poisson negative log-likelihood exp(o) - t o (log t! dropped) on
[-30, 8], continued linearly outside, and their log-variance slot is unused
(no loss, zero gradient).
Used to check the GPU gradients; pinned itself by finite differences.
"""
import numpy as np


def layout(n_in, width, depth, ns):
    in_pad = (n_in + 63) // 64 * 64
    out_pad = (2 * ns + 127) // 128 * 128
    layers, off = [], 0
    for l in range(depth + 1):
        i = in_pad if l == 0 else width
        o = out_pad if l == depth else width
        layers.append((i, o, off, off + i * o))
        off += i * o + o
    return layers, off, in_pad, out_pad


def unpack(theta, layers):
    Ws, bs = [], []
    for (i, o, wo, bo) in layers:
        Ws.append(theta[wo:wo + i * o].reshape(o, i))
        bs.append(theta[bo:bo + o])
    return Ws, bs


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(theta, x, n_in, width, depth, ns):
    layers, _, in_pad, out_pad = layout(n_in, width, depth, ns)
    Ws, bs = unpack(np.asarray(theta, np.float64), layers)
    h = np.zeros((x.shape[0], in_pad))
    h[:, :n_in] = x
    zs, hs = [], [h]
    for l in range(depth):
        z = hs[-1] @ Ws[l].T + bs[l]
        zs.append(z)
        hs.append(silu(z))
    out = hs[-1] @ Ws[depth].T + bs[depth]
    return out, zs, hs


def loss_and_grad(theta, x, t, n_in, width, depth, ns, n_counts=0):
    layers, n_par, in_pad, out_pad = layout(n_in, width, depth, ns)
    theta = np.asarray(theta, np.float64)
    Ws, bs = unpack(theta, layers)
    out, zs, hs = forward(theta, x, n_in, width, depth, ns)
    B = x.shape[0]
    mu = out[:, :ns]
    lv_raw = out[:, ns:2 * ns]
    lv = np.clip(lv_raw, -12.0, 12.0)
    r = t - mu
    inv = np.exp(-lv)
    scale = 1.0 / (B * ns)
    el = 0.5 * (lv + r * r * inv)
    dO = np.zeros_like(out)
    dO[:, :ns] = -r * inv * scale
    dO[:, ns:2 * ns] = np.where((lv_raw < -12.0) | (lv_raw > 12.0), 0.0,
                                0.5 * (1.0 - r * r * inv) * scale)
    c = n_counts
    if c:
        o_raw = mu[:, :c]
        o = np.clip(o_raw, -30.0, 8.0)
        lam = np.exp(o)
        el[:, :c] = (lam - t[:, :c] * o) + (lam - t[:, :c]) * (o_raw - o)
        dO[:, :c] = (lam - t[:, :c]) * scale
        dO[:, ns:ns + c] = 0.0
    loss = float(np.sum(el) * scale)
    grad = np.zeros(n_par)
    d = dO
    for l in range(depth, -1, -1):
        i, o, wo, bo = layers[l]
        grad[wo:wo + i * o] = (d.T @ hs[l]).ravel()
        grad[bo:bo + o] = d.sum(0)
        if l > 0:
            d = (d @ Ws[l]) * dsilu(zs[l - 1])
    return loss, grad
