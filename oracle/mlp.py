"""Surrogate network on the CPU (oracle for the BF16 matrix-core kernels).

TEST INFRASTRUCTURE ONLY (see the package header).  The network has no
reference counterpart: it is the learned part of the fitting loop.  An
ensemble member is a fully connected network x -> W_L gelu(... gelu(W_0 x +
b_0) ...) + b_L with GELU(x) = x Phi(x), trained on the mean squared error.

``bf16=True`` rounds to bfloat16 (round to nearest even) exactly where the GPU
kernels store BF16: the network input, the weights used by the products, the
hidden activations, the output gradient and every hidden pre-activation
gradient.  Products and sums are float64 here (FP32 on the GPU).
"""

import numpy as np
from scipy.special import erf


def to_bf16(x):
    """Round float32 values to bfloat16 (returned as float64)."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r.astype(np.uint32)).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def unpack(params, dims):
    """[(W_l [out][in], b_l [out])] from one member's flat parameter vector."""
    out, off = [], 0
    for l in range(len(dims) - 1):
        i, o = dims[l], dims[l + 1]
        W = params[off: off + o * i].reshape(o, i)
        off += o * i
        b = params[off: off + o]
        off += o
        out.append((np.asarray(W, np.float64), np.asarray(b, np.float64)))
    return out


def _q(x, bf16):
    return to_bf16(x) if bf16 else np.asarray(x, np.float64)


def forward(params, dims, x, bf16=False):
    """Output [B][dims[-1]] and the cache for ``backward``."""
    layers = unpack(params, dims)
    h = _q(x, bf16)
    cache = []
    for l, (W, b) in enumerate(layers):
        z = h @ _q(W, bf16).T + b
        cache.append((h, z))
        if l + 1 < len(layers):
            h = _q(gelu(z), bf16)
        else:
            h = z
    return h, cache


def loss_and_grad(params, dims, x, t, bf16=False):
    """Mean squared error over all outputs and its gradient (flat, like params)."""
    y, cache = forward(params, dims, x, bf16)
    B, N = y.shape
    d = y - np.asarray(t, np.float64)
    loss = float(np.mean(d * d))
    layers = unpack(params, dims)
    grads = []
    dz = _q(2.0 * d / (B * N), bf16)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        h_in, _ = cache[l]
        gW = dz.T @ h_in
        gb = dz.sum(axis=0)
        grads.append((gW, gb))
        if l > 0:
            dh = dz @ _q(W, bf16)
            dz = _q(dh * gelu_grad(cache[l - 1][1]), bf16)
    flat = []
    for gW, gb in grads[::-1]:
        flat.append(gW.reshape(-1))
        flat.append(gb)
    g = np.concatenate(flat)
    out = np.zeros(len(params))
    out[: len(g)] = g
    return loss, out
