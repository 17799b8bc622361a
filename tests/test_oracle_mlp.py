"""Pins of the surrogate network's CPU oracle (oracle/mlp_oracle.py) by
mathematics: finite differences, closed forms, hand-worked bf16 bit patterns
and the published Adam update."""
import math

import numpy as np
import pytest

import mlp_oracle as mo


def _tiny(seed=0, dims=(7, 5, 6, 3)):
    rng = np.random.default_rng(seed)
    layers = [(rng.normal(size=(dout, din)) / math.sqrt(din), rng.normal(size=dout) * 0.1)
              for din, dout in zip(dims[:-1], dims[1:])]
    X = rng.normal(size=(4, dims[0]))
    T = rng.normal(size=(4, dims[-1]))
    return layers, X, T


def test_backward_matches_finite_differences():
    layers, X, T = _tiny()
    L, grads = mo.backward(layers, X, T, 3, 2)  # padding: last row and last output ignored
    h = 1e-6
    worst = 0.0
    for l, (W, b) in enumerate(layers):
        for arr, g in ((W, grads[l][0]), (b, grads[l][1])):
            it = np.nditer(arr, flags=["multi_index"])
            for _ in it:
                idx = it.multi_index
                old = arr[idx]
                arr[idx] = old + h
                lp = mo.loss(layers, X, T, 3, 2)
                arr[idx] = old - h
                lm = mo.loss(layers, X, T, 3, 2)
                arr[idx] = old
                fd = (lp - lm) / (2 * h)
                worst = max(worst, abs(fd - g[idx]) / max(1e-6, abs(fd)))
    assert abs(L - mo.loss(layers, X, T, 3, 2)) < 1e-15
    assert worst < 1e-5


def test_padding_rows_and_columns_do_not_contribute():
    layers, X, T = _tiny(1)
    T2 = T.copy()
    T2[3, :] = 1e3
    T2[:, 2] = -1e3
    a = mo.backward(layers, X, T, 3, 2)
    b = mo.backward(layers, X, T2, 3, 2)
    assert a[0] == b[0]
    for (ga, gb_), (ha, hb) in zip(a[1], b[1]):
        assert np.array_equal(ga, ha) and np.array_equal(gb_, hb)


def test_gelu_closed_forms():
    z = np.array([0.0, 1.0, -1.0, 8.0, -8.0])
    g = mo.gelu(z)
    assert g[0] == 0.0
    # gelu(1) = Phi(1) = 0.8413447460685429...
    assert abs(g[1] - 0.8413447460685429) < 1e-15
    assert abs(g[2] - (-1 + 0.8413447460685429)) < 1e-15
    assert abs(g[3] - 8.0) < 1e-12 and abs(g[4]) < 1e-12
    assert mo.gelu_grad(np.array(0.0)) == 0.5
    zz = np.linspace(-4, 4, 33)
    fd = (mo.gelu(zz + 1e-6) - mo.gelu(zz - 1e-6)) / 2e-6
    assert np.max(np.abs(fd - mo.gelu_grad(zz))) < 1e-8


def test_bf16_rounding_bit_patterns():
    def bits(x):
        return mo.bf16(np.array([x], dtype=np.float32)).view(np.uint32)[0] >> 16

    assert bits(1.0) == 0x3F80
    assert bits(-2.0) == 0xC000
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1.0
    assert bits(1.0 + 2.0 ** -8) == 0x3F80
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 (odd) and 1 + 2^-6: -> up
    assert bits(1.0 + 3 * 2.0 ** -8) == 0x3F82
    # just above the halfway point rounds up
    assert bits(np.nextafter(np.float32(1.0 + 2.0 ** -8), np.float32(2))) == 0x3F81
    assert bits(np.float32(np.inf)) == 0x7F80
    nan = mo.bf16(np.array([np.nan], dtype=np.float32))[0]
    assert np.isnan(nan)
    # the largest float32 rounds to infinity in bf16
    assert bits(np.finfo(np.float32).max) == 0x7F80


def test_adam_matches_published_update_by_hand():
    w, g = 0.5, 0.2
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    w1, m1, v1 = mo.adam(w, g, 0.0, 0.0, 1, lr, b1, b2, eps)
    # step 1: m_hat = g, v_hat = g^2 -> update = lr * g / (|g| + eps)
    assert abs(w1 - (w - lr * g / (abs(g) + eps))) < 1e-15
    w2, m2, v2 = mo.adam(w1, -0.1, m1, v1, 2, lr, b1, b2, eps)
    m_ref = 0.9 * 0.02 + 0.1 * -0.1
    v_ref = 0.999 * 0.001 * 0.04 + 0.001 * 0.01
    upd = lr * (m_ref / (1 - 0.81)) / (math.sqrt(v_ref / (1 - 0.999 ** 2)) + eps)
    assert abs(w2 - (w1 - upd)) < 1e-15


def test_emulated_passes_stay_close_to_exact():
    """The bf16 emulation moves loss and gradients by bf16-sized amounts only."""
    rng = np.random.default_rng(3)
    dims = [64, 64, 64, 64]
    layers = [(rng.normal(size=(o, i)) / 8, rng.normal(size=o) * 0.1) for i, o in zip(dims[:-1], dims[1:])]
    X = rng.normal(size=(64, 64))
    T = rng.normal(size=(64, 64))
    L0, g0 = mo.backward(layers, X, T, 64, 64)
    L1, g1 = mo.backward(layers, X, T, 64, 64, emulate=True)
    assert abs(L1 - L0) / L0 < 2e-2
    a = np.concatenate([g[0].ravel() for g in g0])
    b = np.concatenate([g[0].ravel() for g in g1])
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 5e-2


def test_device_normal_cdf_coefficients_reproduce_erfc():
    """The GEMM epilogues evaluate GELU as z * Phi(z) with erfc from the
    Numerical Recipes Chebyshev fit (erfcc); the coefficients are read from
    the kernel source and checked against math.erfc: fractional error below
    1.2e-7 (the published bound) on [0, 9], and Phi(z) + Phi(-z) = 1."""
    import os
    import re
    src = open(os.path.join(os.path.dirname(__file__), "..", "ripple.jl_amd", "csrc", "mlp.hip")).read()
    body = src[src.index("float phi_cdf(float z)"):src.index("float gelu(float z)")]
    first = float(re.search(r"float p = ([-0-9.e]+)f;", body).group(1))
    rest = [float(c) for c in re.findall(r"p = fmaf\(p, t, ([-0-9.e]+)f\);", body)]
    assert len(rest) == 9
    coef = [first] + rest

    def erfc(x):
        t = 1.0 / (1.0 + 0.5 * x)
        p = coef[0]
        for c in coef[1:]:
            p = p * t + c
        return t * math.exp(-x * x + p)

    xs = np.linspace(0.0, 9.0, 9001)
    worst = max(abs(erfc(x) - math.erfc(x)) / math.erfc(x) for x in xs)
    assert worst < 1.2e-7, worst
    for z in (0.3, 1.7, 4.2):
        phi = lambda v: 1.0 - 0.5 * erfc(v / math.sqrt(2)) if v >= 0 else 0.5 * erfc(-v / math.sqrt(2))
        assert abs(phi(z) + phi(-z) - 1.0) < 1e-15
        assert abs(z * phi(z) - mo.gelu(np.array(z))) < 1.2e-7 * abs(z)
