"""CPU restatement of the learned surrogate's network.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module; the product package never does.

The surrogate has no counterpart in the reference (it stands in for the
per-column source evaluation of Physics/Packages/precomputed_sources.jl:240-261),
so its oracle is pinned by mathematics instead of reference vectors:
  * `backward` in float64 without rounding is checked against central finite
    differences of `loss`;
  * GELU and its derivative against closed-form values (gelu(0) = 0,
    gelu'(0) = 1/2, tails) and against a finite difference of `gelu`;
  * `bf16` against hand-worked bit patterns (ties to even, NaN, overflow);
  * `adam` against the published update (Kingma and Ba, 2015, Algorithm 1)
    written out for one scalar.
With `emulate=True` the forward and backward passes round exactly where the
device kernels store bf16 (inputs, weights, activations, dY, dZ), so the
remaining difference to the device is FP32 vs FP64 accumulation order.

Layout: parameters are the concatenation over layers of W_l [dout x din]
(row-major) and b_l [dout], on padded dims (multiples of 64).  Loss =
mean over valid rows and valid output columns of (Y - T)^2.
"""
import math

import numpy as np
from scipy.special import erf


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def unpack(flat, dims):
    out, off = [], 0
    for din, dout in zip(dims[:-1], dims[1:]):
        W = np.asarray(flat[off:off + din * dout], dtype=np.float64).reshape(dout, din)
        off += din * dout
        b = np.asarray(flat[off:off + dout], dtype=np.float64)
        off += dout
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def _r(x, emulate):
    return bf16(x).astype(np.float64) if emulate else np.asarray(x, dtype=np.float64)


def pad_rows(X, width, rows):
    out = np.zeros((rows, width))
    out[:X.shape[0], :X.shape[1]] = X
    return out


def forward(layers, X, emulate=False):
    """Returns (Y, cache); X already padded to the input width."""
    a = _r(X, emulate)
    cache = [a]
    zs = []
    for l, (W, b) in enumerate(layers):
        z = a @ _r(W, emulate).T
        if emulate:
            z = z.astype(np.float32).astype(np.float64)  # FP32 accumulator
        z = z + b
        if l < len(layers) - 1:
            if emulate:
                z = z.astype(np.float32).astype(np.float64)
            zs.append(z)
            a = _r(gelu(z), emulate)
            cache.append(a)
        else:
            Y = z
    return Y, (cache, zs)


def loss(layers, X, T, valid_rows, valid_cols, emulate=False):
    Y, _ = forward(layers, X, emulate)
    d = (Y - T)[:valid_rows, :valid_cols]
    return float(np.mean(d * d))


def backward(layers, X, T, valid_rows, valid_cols, emulate=False):
    """(loss, [(gW, gb)]) of the mean squared error."""
    Y, (acts, zs) = forward(layers, X, emulate)
    d = np.zeros_like(Y)
    d[:valid_rows, :valid_cols] = (Y - T)[:valid_rows, :valid_cols]
    count = valid_rows * valid_cols
    L = float(np.sum(d * d) / count)
    dZ = _r(2.0 * d / count, emulate)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        gW = dZ.T @ acts[l]
        gb = dZ.sum(axis=0)
        grads[l] = (gW, gb)
        if l > 0:
            dA = dZ @ _r(W, emulate)
            dZ = _r(dA * gelu_grad(zs[l - 1]), emulate)
    return L, grads


def adam(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam update; returns (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
