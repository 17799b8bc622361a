"""Pins the oracle's DiscreteInteractionApproximation{ListedQuadruplets}
(Sources/Sources.jl:1332-1385) against the reference's own known answer
(test/sources/source_terms.jl:1098-1161) and a hand-derived closed form for
a transfer power other than one (CPU only)."""
import ctypes
import math

import numpy as np
import pytest


def _hook_tendency(oracle, N, interactions, power, w):
    nx, ny, nk, nphi = N.shape
    hook, user, ctx = oracle.listed_hook(N.shape, interactions, power, w)
    Nf = oracle._c(N)
    G = np.zeros_like(Nf)
    fn = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)(hook.value)
    fn(Nf.ctypes.data, G.ctypes.data, user)
    return G.reshape(N.shape, order="F")


def _weights(kf, pf):
    return np.array([[(kf[m + 1] ** 2 - kf[m] ** 2) / 2 * (pf[n + 1] - pf[n]) for n in range(len(pf) - 1)]
                     for m in range(len(kf) - 1)])


def test_reference_known_answer(oracle):
    """source_terms.jl:1098-1161: frequency 0.2 Hz, four directions, donors
    (1,1) and (1,3) with N = 4 and 3, receivers (1,2) and (1,4), rate 0.05,
    power 1: flux 0.6 leaves each donor and enters each receiver, and the
    weighted action, energy and wavevector tendencies vanish."""
    g = 9.81
    ff = np.array([0.0, 0.4])  # lower-bounded faces of the single centre 0.2
    kf = (2 * math.pi * ff) ** 2 / g
    phi = np.array([0.0, math.pi / 2, math.pi, 3 * math.pi / 2])
    pf = np.concatenate([phi - math.pi / 4, [phi[-1] + math.pi / 4]])
    w = _weights(kf, pf)
    N = np.full((1, 1, 1, 4), 0.5, order="F")
    N[0, 0, 0, :] = [4.0, 0.25, 3.0, 0.75]
    quad = [((0, 0), (0, 2), (0, 1), (0, 3), 0.05)]
    G = _hook_tendency(oracle, N, quad, 1.0, w)[0, 0, 0]
    flux = 0.05 * 4.0 * 3.0
    np.testing.assert_allclose(G, [-flux / w[0, 0], flux / w[0, 1], -flux / w[0, 2], flux / w[0, 3]],
                               rtol=1e-15, atol=0)
    kappa = (2 * math.pi * 0.2) ** 2 / g
    assert abs((G * w[0]).sum()) < 1e-14
    assert abs((G * w[0] * 0.2).sum()) < 1e-14
    assert abs((G * w[0] * kappa * np.cos(phi)).sum()) < 1e-14
    assert abs((G * w[0] * kappa * np.sin(phi)).sum()) < 1e-14


def test_power_and_clipping(oracle):
    """Closed form with power 1.5, two interactions sharing a receiver and a
    negative donor that switches its interaction off."""
    rng = np.random.default_rng(2)
    w = rng.uniform(0.5, 2.0, (3, 4))
    N = np.zeros((2, 1, 3, 4), order="F")
    N[..., :, :] = rng.uniform(0.1, 1.0, (2, 1, 3, 4))
    N[1, 0, 2, 0] = -0.3  # donor of the second interaction in column 1
    quads = [((0, 0), (0, 2), (0, 1), (0, 3), 0.2), ((2, 0), (2, 2), (2, 1), (0, 3), 0.7)]
    p = 1.5
    G = _hook_tendency(oracle, N, quads, p, w)
    for i in range(2):
        a = N[i, 0]
        want = np.zeros((3, 4))
        for (d1, d2, r1, r2, rate) in quads:
            flux = rate * max(a[d1], 0) ** p * max(a[d2], 0) ** p
            if flux == 0:
                continue
            for r in (r1, r2):
                want[r] += flux / w[r]
            for d in (d1, d2):
                if a[d] > 0:
                    want[d] -= flux / w[d]
        np.testing.assert_allclose(G[i, 0], want, rtol=1e-14, atol=1e-300)
    assert G[1, 0, 2, 1] == 0  # column 1 lost the second interaction


def test_rate_field_is_each_columns_constant_rate(oracle):
    """Rates given per column (the reference's source_value of an array or a
    function of position, Sources.jl:1450-1467): every column's tendency is
    the constant-rate tendency of that column alone with its own rates, a
    field equal to the constants everywhere changes nothing, and the
    tendency is linear in each column's rate."""
    rng = np.random.default_rng(4)
    nx, ny, nk, nphi = 3, 2, 3, 4
    w = rng.uniform(0.5, 2.0, (nk, nphi))
    N = np.asfortranarray(rng.uniform(0.1, 1.0, (nx, ny, nk, nphi)))
    quads = [((0, 0), (0, 2), (0, 1), (0, 3), 0.2), ((2, 0), (2, 2), (2, 1), (0, 3), 0.7)]
    field = rng.uniform(0.0, 2.0, (nx, ny, len(quads)))
    field[1, 0, 0] = 0.0  # a column where the first interaction is off
    G = _hook_tendency_field(oracle, N, quads, 1.5, w, field)
    for i in range(nx):
        for j in range(ny):
            own = [q[:4] + (field[i, j, k],) for k, q in enumerate(quads)]
            want = _hook_tendency(oracle, N[i:i + 1, j:j + 1], own, 1.5, w)[0, 0]
            np.testing.assert_array_equal(G[i, j], want)
    const = np.broadcast_to(np.array([q[4] for q in quads]), (nx, ny, len(quads)))
    np.testing.assert_array_equal(_hook_tendency_field(oracle, N, quads, 1.5, w, const),
                                  _hook_tendency(oracle, N, quads, 1.5, w))
    np.testing.assert_allclose(_hook_tendency_field(oracle, N, quads, 1.5, w, 2 * field), 2 * G, rtol=1e-14,
                               atol=1e-300)


def _hook_tendency_field(oracle, N, interactions, power, w, field):
    hook, user, ctx = oracle.listed_hook(N.shape, interactions, power, w, rate_field=field)
    Nf = oracle._c(N)
    G = np.zeros_like(Nf)
    fn = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)(hook.value)
    fn(Nf.ctypes.data, G.ctypes.data, user)
    return G.reshape(N.shape, order="F")
