"""Pins the oracle's DiscreteInteractionApproximation{SymmetricQuadruplet}
(Physics/NonlinearInteractions/symmetric_quadruplet.jl) against the
reference's own tests (test/physics/symmetric_quadruplet.jl), exact
invariants (quadratic homogeneity, rotation equivariance) and hand-derived
closed forms for a donor whose receivers sit on grid rows (CPU only)."""
import ctypes
import math

import numpy as np
import pytest

G = 9.81


def _jonswap_like(nk=16, nphi=24):
    f = np.geomspace(0.05, 0.4, nk)
    kappa = (2 * math.pi * f) ** 2 / G
    phi = np.linspace(0, 2 * math.pi, nphi + 1)[:nphi]
    N = np.zeros((1, 1, nk, nphi), order="F")
    for n in range(nphi):
        for m in range(nk):
            ff = math.sqrt(G * kappa[m]) / (2 * math.pi)
            N[0, 0, m, n] = 1e-4 * math.exp(-(ff - 0.15) ** 2 / 0.02 ** 2) * max(math.cos(phi[n]), 0) ** 2
    return N, kappa, phi


def test_reference_known_answers(oracle):
    # test/physics/symmetric_quadruplet.jl:1-17: offsets 11.48 and 33.56 degrees at lambda = 0.25.
    dp, dm = oracle.dia_offsets(0.25)
    assert dp == pytest.approx(math.radians(11.48), abs=0.01)
    assert dm == pytest.approx(math.radians(33.56), abs=0.01)
    # :20-30: zero spectrum, zero transfer.
    N, kappa, phi = _jonswap_like(12, 16)
    assert not oracle.dia_transfer(np.zeros_like(N), kappa, phi).any()
    # :32-61: the spectral-measure-weighted total is bounded by the donor magnitude.
    N, kappa, phi = _jonswap_like(16, 24)
    T = oracle.dia_transfer(N, kappa, phi)
    kf = oracle.centers_to_faces(kappa, lower_bounded=True, lower=0.0)
    w = oracle.geometry(kappa, kf, phi, oracle.phi_faces(phi))["w"]
    total = (T[0, 0] * w).sum()
    magnitude = (np.abs(T[0, 0]) * w).sum()
    assert magnitude > 0 and abs(total) < 1.5 * magnitude


def test_invariants(oracle):
    rng = np.random.default_rng(0)
    N, kappa, phi = _jonswap_like(16, 24)
    N = np.asfortranarray(np.broadcast_to(N, (2, 3) + N.shape[2:]) * rng.uniform(0.5, 1.5, (2, 3, 16, 24)))
    T = oracle.dia_transfer(N, kappa, phi)
    # Every term is a product of two energies: doubling N quadruples T exactly.
    assert np.array_equal(oracle.dia_transfer(2 * N, kappa, phi), 4 * T)
    # Rotating the spectrum by one direction bin rotates the transfer.
    Tr = oracle.dia_transfer(np.roll(N, 1, axis=3), kappa, phi)
    assert np.abs(Tr - np.roll(T, 1, axis=3)).max() <= 1e-12 * np.abs(T).max()
    # Columns are independent.
    T1 = oracle.dia_transfer(np.asfortranarray(N[1:2, 2:3]), kappa, phi)
    assert np.array_equal(T1[0, 0], T[1, 2])


def _aligned_grid(lam=0.21, nk=8, nphi=72):
    """sigma_m = sigma_0 1.1^m, so (1 + lambda) sigma lands two rows up for lambda = 0.21."""
    s0 = 2 * math.pi * 0.08
    sigma = s0 * 1.1 ** np.arange(nk)
    kappa = sigma ** 2 / G
    phi = np.arange(nphi) * (2 * math.pi / nphi)
    return sigma, kappa, phi


def _offsets(lam):
    a, b = (1 + lam) ** 2, (1 - lam) ** 2
    return math.acos((a * a - b * b + 4) / (4 * a)), math.acos((b * b - a * a + 4) / (4 * b))


@pytest.mark.parametrize("side", ["plus", "minus"])
def test_closed_form_single_quadruplet(oracle, side):
    """A donor at (row 3, 50 deg) with action A, and action B on the four
    corner cells of one of its receivers for orientation +1; no other
    quadruplet has both partners occupied (checked by the geometry of a
    5-degree grid), so T is the single term
        T_E = C/g^4 f_a^11 E_a E_r DAL / 2
    taken twice from the donor (both receivers in range) and spread to the
    receiver corners with their bilinear weights."""
    lam, C = 0.21, 2.78e7
    sigma, kappa, phi = _aligned_grid(lam)
    nk, nphi = len(kappa), len(phi)
    dphi = 2 * math.pi / nphi
    dtp, dtm = _offsets(lam)
    ma, na, A, B = 3, 10, 2.0e-3, 5.0e-4
    Ef = 4 * math.pi * sigma ** 2 / G

    def corners(rows, theta):
        nr = theta / dphi
        cols = [(int(nr) % nphi, 1 - (nr - int(nr))), ((int(nr) + 1) % nphi, nr - int(nr))]
        return [(r, c, wr * wc) for r, wr in rows for c, wc in cols]

    mr = math.log((1 - lam) * sigma[ma] / sigma[0]) / math.log(1.1)  # fractional row of (1 - lambda) sigma
    plus = corners([(5, 1.0)], phi[na] + dtp)  # (1 + lambda) sigma_3 = sigma_5
    minus = corners([(int(mr), 1 - (mr - int(mr))), (int(mr) + 1, mr - int(mr))], phi[na] - dtm)
    occupied, dal = (plus, 1 / (1 + lam) ** 4) if side == "plus" else (minus, 1 / (1 - lam) ** 4)
    N = np.zeros((1, 1, nk, nphi), order="F")
    N[0, 0, ma, na] = A
    for r, c, _ in occupied:
        N[0, 0, r, c] = B
    T = oracle.dia_transfer(N, kappa, phi, coefficient=C, frequency_offset=lam, gravity=G)[0, 0]
    Er = sum(w * Ef[r] * B for r, c, w in occupied)
    fa = sigma[ma] / (2 * math.pi)
    TE = C / G ** 4 * fa ** 11 * (Ef[ma] * A) * Er * dal / 2
    # The donor gives twice (both receivers in range); both receivers' corners gain.
    want = np.zeros((nk, nphi))
    want[ma, na] = -2 * TE / Ef[ma]
    for r, c, w in plus + minus:
        want[r, c] += w * TE / Ef[r]
    assert np.abs(T).max() > 0
    np.testing.assert_allclose(T, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_bundle_split(oracle):
    """precomputed_sources.jl:240-283 with the nonlinear member alone: the
    tendency is T where T >= 0, else -min(-T / max(N, eps), 1) N."""
    N, kappa, phi = _jonswap_like(12, 16)
    T = oracle.dia_transfer(N, kappa, phi)
    src = oracle.OrcSources()
    src.gravity = 9.81
    src.nonlinear = 1
    src.dia_coefficient, src.dia_frequency_offset, src.dia_gravity = 2.78e7, 0.25, 9.81
    kf = oracle.centers_to_faces(kappa, lower_bounded=True, lower=0.0)
    pf = oracle.phi_faces(phi)
    w = oracle._c(oracle.geometry(kappa, kf, phi, pf)["w"])
    xn = yn = np.array([0.5])
    kc, pc, kfc, pfc = (oracle._c(a) for a in (kappa, phi, kf, pf))
    ctx = oracle.OrcSourceCtx(ctypes.pointer(src), 1, 1, len(kappa), len(phi), xn.ctypes.data, yn.ctypes.data,
                              kc.ctypes.data, pc.ctypes.data, w.ctypes.data, 0.0, kfc.ctypes.data, pfc.ctypes.data)
    Nf = oracle._c(N)
    Gt = np.zeros_like(Nf)
    oracle.lib().orc_sources_hook_f64(oracle._p(Nf), oracle._p(Gt), ctypes.pointer(ctx))
    Gt = Gt.reshape(N.shape, order="F")
    eps = np.finfo(np.float64).eps
    want = np.where(T >= 0, T, -np.minimum(-T / np.maximum(N, eps), 1.0) * N)
    np.testing.assert_allclose(Gt, want, rtol=1e-15, atol=0)
    assert (T < 0).any() and (T > 0).any()
