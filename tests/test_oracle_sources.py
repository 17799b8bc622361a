"""Pins the oracle's source terms and winds (CPU only): the reference's known
answers (test/physics/shared.jl, test/physics/st3.jl, test/forcing/winds.jl)
and closed forms on single-bin spectra."""
import math

import numpy as np
import pytest

from ripple_amd import grids, workloads
import ripple_amd as ra
from ripple_amd.sources import to_struct


def src_struct(oracle, sources):
    return oracle.sources_from(to_struct(sources))


def tiny(nfreq=4, nphi=8, nx=3, ny=2, N=1e-3, polar=False):
    g = grids.RectilinearGrid(size=(nx, ny, 1), x=(0, 3.0), y=(0, 2.0), z=(0, 1))
    if polar:
        sg = grids.PolarWaveVectorGrid(kappa=np.linspace(0.05, 0.4, nfreq), phi=np.linspace(0, 2 * math.pi, nphi + 1)[:nphi])
    else:
        sg = grids.FrequencyDirectionGrid(frequency=np.linspace(0.1, 0.3, nfreq),
                                          phi=np.linspace(0, 2 * math.pi * (nphi - 1) / nphi, nphi))
    N0 = np.full((nx, ny, nfreq, nphi), N, order="F")
    return workloads.Workload("tiny", g, sg, None, N0, 1.0, 1, dict(horizontal_advection=None, timestepper="ForwardEuler"))


def test_drag_known_answers():
    d = ra.BulkWindDrag("linear")
    assert d.friction_velocity(10.0) == pytest.approx(10.0 * math.sqrt(1.45e-3), rel=1e-12)
    c = ra.BulkWindDrag("capped")
    assert c.friction_velocity(50.0) > c.friction_velocity(30.0) > c.friction_velocity(20.0)


def test_holland_wind_known_answers(oracle):
    """test/forcing/winds.jl:34-46: speed = vmax at r = rmax, 0 at the centre,
    angle pi/2 east of a counter-clockwise storm, -pi/2 when clockwise."""
    track = ra.LinearStormTrack([0.0, 1.0], [(0.0, 0.0), (1.0, 0.0)])
    w = ra.HollandHurricaneWind(center=track, vmax=0.3, rmax=1.0, radius=4.0, shape_parameter=1.5)
    s = src_struct(oracle, ra.PressureCorrelationInput(drag=ra.BulkWindDrag(), wind=w))
    U, th, us = oracle.wind_columns(s, [1.0, 2.0, 0.0, 3.0], [0.0], 0.0)
    assert U[0, 0] == pytest.approx(0.3) and th[0, 0] == pytest.approx(math.pi / 2)
    assert U[2, 0] == 0.0 and U[3, 0] < 0.3
    U1, _, _ = oracle.wind_columns(s, [2.0], [0.0], 1.0)  # the storm moved to x = 1
    assert U1[0, 0] == pytest.approx(0.3)
    cw = ra.HollandHurricaneWind(vmax=0.3, rmax=1.0, radius=4.0, rotation="clockwise")
    s = src_struct(oracle, ra.PressureCorrelationInput(drag=ra.BulkWindDrag(), wind=cw))
    _, th, _ = oracle.wind_columns(s, [1.0], [0.0], 0.0)
    assert th[0, 0] == pytest.approx(-math.pi / 2)
    # Beyond `radius` the profile decays exponentially with scale radius - rmax.
    U5, _, _ = oracle.wind_columns(s, [5.0, 6.0], [0.0], 0.0)
    st = lambda r: 0.3 * math.sqrt((1 / r) ** 1.5 * math.exp(1 - (1 / r) ** 1.5))
    assert U5[0, 0] == pytest.approx(st(5.0) * math.exp(-1 / 3), rel=1e-13)
    assert U5[1, 0] == pytest.approx(st(6.0) * math.exp(-2 / 3), rel=1e-13)


def test_wind_input_growth_signs(oracle):
    """test/physics/st3.jl:18-40: with 15 m/s some bin grows and none decays;
    calm wind gives exactly zero."""
    wl = tiny()
    for wind, check in ((15.0, "grow"), (0.0, "calm")):
        src = src_struct(oracle, ra.PressureCorrelationInput(drag=ra.BulkWindDrag(), wind=wind))
        om = oracle.OracleSourceModel(wl, src, transport=oracle.MODE_NONE)
        G = om.tendency()
        if check == "grow":
            assert G.max() > 0 and G.min() >= 0
            # Waves against the wind (cos(theta - theta_u) <= 0) never grow.
            phi = wl.spectral_grid.phi
            assert np.all(G[:, :, :, np.cos(phi) <= 1e-12] == 0)
        else:
            assert np.all(G == 0)


def _jp_rates(oracle, U10, kappa, phi, N=2e-3, **params):
    """Per-cell growth rate G / N of PressureCorrelationInput alone on one
    column of a polar grid, wind along phi = 0."""
    g = grids.RectilinearGrid(size=(1, 1, 1), x=(0, 1.0), y=(0, 1.0), z=(0, 1))
    sg = grids.PolarWaveVectorGrid(kappa=list(kappa), phi=list(phi))
    N0 = np.full((1, 1, len(kappa), len(phi)), N, order="F")
    wl = workloads.Workload("jp", g, sg, None, N0, 1.0, 1, dict(horizontal_advection=None, timestepper="ForwardEuler"))
    src = src_struct(oracle, ra.PressureCorrelationInput(drag=ra.BulkWindDrag(), wind=U10, **params))
    return oracle.OracleSourceModel(wl, src, transport=oracle.MODE_NONE).tendency()[0, 0] / N


def test_jp_rate_published_form(oracle):
    """Janssen (1991) / Bidlot et al. (2005) quasi-linear growth in its
    published form: gamma = eps (beta_max / kappa_vk^2) mu ln^4(mu) x^2 sigma
    for mu = k z0 exp(kappa_vk / x) < 1, with x = (u*/c + z_alpha) cos(dtheta),
    Charnock z0 = alpha0 u*^2 / g, eps = rho_a / rho_w, and the reference's
    direction factor cos^p_in of which x^2 holds cos^2
    (pressure_correlation.jl:109-136).  The product-and-logarithm form here
    is not the oracle's exponent-of-a-sum form, so a slip in either differs."""
    phi = list(np.arange(9) * (2 * math.pi / 9))
    for U10, kappa in ((20.0, [0.05, 0.1, 0.3, 1.0]), (8.0, [0.2, 0.6, 2.0])):
        R = _jp_rates(oracle, U10, kappa, phi)
        g, kvk = 9.81, 0.4
        us = U10 * math.sqrt((0.8 + 0.065 * U10) * 1e-3)  # BulkWindDrag(:linear), Cd(10) = 1.45e-3
        z0 = 0.0095 * us ** 2 / g
        grew = 0
        for m, k in enumerate(kappa):
            c = math.sqrt(g / k)
            for n, th in enumerate(phi):
                cs = math.cos(th)
                if cs <= 0:
                    assert R[m, n] == 0.0
                    continue
                x = (us / c + 0.011) * cs
                mu = k * z0 * math.exp(kvk / x)
                if mu >= 1:
                    assert R[m, n] == 0.0
                    continue
                gamma = (1.225 / 1025.0) * (1.2 / kvk ** 2) * mu * math.log(mu) ** 4 * x ** 2 * math.sqrt(g * k)
                assert R[m, n] == pytest.approx(gamma, rel=1e-12)
                grew += 1
        assert grew > 0


def test_jp_rate_scalings(oracle):
    """Exact consequences of the rate's structure: linear in beta_max and in
    rho_air / rho_water; raising p_in by one multiplies every growing bin by
    its cos(theta - theta_u); the growth window depends on none of them."""
    kappa, phi = [0.05, 0.1, 0.3], list(np.arange(16) * (2 * math.pi / 16))
    base = _jp_rates(oracle, 15.0, kappa, phi)
    assert np.count_nonzero(base) >= 6
    np.testing.assert_allclose(_jp_rates(oracle, 15.0, kappa, phi, beta_max=2.4), 2 * base, rtol=1e-14)
    np.testing.assert_allclose(_jp_rates(oracle, 15.0, kappa, phi, rho_air=2.45), 2 * base, rtol=1e-14)
    p3 = _jp_rates(oracle, 15.0, kappa, phi, p_in=3.0)
    np.testing.assert_allclose(p3, base * np.cos(phi)[None, :], rtol=1e-13)


def test_jp_rate_vanishes_quartically_at_the_window_edge(oracle):
    """Inside the window the rate carries ln^4(mu) with ln(mu) linear in the
    distance to the edge, so just inside the edge halving the distance divides
    the rate by 2^4.  The edge wind is found by bisection on the oracle's
    own output (zero outside, positive inside)."""
    kappa, phi = [0.05], [0.0]
    lo, hi = 5.0, 40.0  # zero at 5 m/s (waves outrun the wind), positive at 40 m/s
    assert _jp_rates(oracle, lo, kappa, phi)[0, 0] == 0.0 and _jp_rates(oracle, hi, kappa, phi)[0, 0] > 0
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if _jp_rates(oracle, mid, kappa, phi)[0, 0] > 0:
            hi = mid
        else:
            lo = mid
    edge = hi
    d = 1e-5 * edge
    r1, r2 = _jp_rates(oracle, edge + d, kappa, phi)[0, 0], _jp_rates(oracle, edge + 2 * d, kappa, phi)[0, 0]
    assert r2 / r1 == pytest.approx(16.0, rel=1e-3)


def test_mean_spectrum_single_bin_closed_form(oracle):
    """With action in one bin only, m0 = N w, k_bar = k, sigma_bar = sqrt(g k), and
    the damping is -C_ds (m0 k^2)^2 sigma (delta1 + delta2) N (mean_spectrum.jl:8-16)."""
    wl = tiny(nfreq=4, nphi=8, nx=1, ny=1)
    N = np.zeros((1, 1, 4, 8), order="F")
    N[0, 0, 2, 3] = 0.7
    wl.N0 = N
    d = ra.MeanSpectrumWhitecapping()
    src = src_struct(oracle, ra.PrecomputedSources(dissipation=d))
    G = oracle.OracleSourceModel(wl, src, transport=oracle.MODE_NONE).tendency()
    sg = wl.spectral_grid
    k = sg.kappa[2]
    m0 = 0.7 * sg.weights[2, 3]
    lam = 2.1 * (m0 * k * k) ** 2 * math.sqrt(9.81 * k) * (0.4 + 0.6)
    assert G[0, 0, 2, 3] == pytest.approx(-lam * 0.7, rel=1e-12)
    assert np.count_nonzero(G) == 1


def test_local_saturation_closed_form(oracle):
    wl = tiny(nfreq=3, nphi=4, nx=1, ny=1, polar=True)
    N = np.zeros((1, 1, 3, 4), order="F")
    N[0, 0, 1, 0], N[0, 0, 2, 2] = 5.0, 1e-9  # one saturated cell, one below threshold
    wl.N0 = N
    d = ra.LocalSaturationDissipation(B_r=1e-2)
    G = oracle.OracleSourceModel(wl, src_struct(oracle, d), transport=oracle.MODE_NONE).tendency()
    k = wl.spectral_grid.kappa[1]
    s = math.sqrt(9.81 * k)
    over = 5.0 * s * k ** 3 / 1e-2 - 1
    assert G[0, 0, 1, 0] == pytest.approx(-2.2e-5 * s * over ** 2 * 5.0, rel=1e-12)
    assert G[0, 0, 2, 2] == 0.0


def test_stress_cap_limits_wave_supported_stress(oracle):
    """The bundle cap scales the growth so that the wave-supported stress along
    the wind is at most 0.4 rho_a u*^2 (precomputed_sources.jl:163-166)."""
    wl = tiny(nfreq=6, nphi=8, nx=1, ny=1)
    wl.N0[...] = 50.0  # a strong sea: the cap is active
    inp = ra.PressureCorrelationInput(drag=ra.BulkWindDrag(), wind=25.0)
    capped = oracle.OracleSourceModel(wl, src_struct(oracle, ra.PrecomputedSources(wind_input=inp)),
                                      transport=oracle.MODE_NONE).tendency()
    raw = oracle.OracleSourceModel(wl, src_struct(oracle, inp), transport=oracle.MODE_NONE).tendency()
    assert np.all(capped <= raw) and capped.max() > 0
    factor = capped[raw > 0] / raw[raw > 0]
    np.testing.assert_allclose(factor, factor.flat[0], rtol=1e-12)  # one factor per column
    sg = wl.spectral_grid
    us = 25.0 * math.sqrt((0.8 + 0.065 * 25.0) * 1e-3)
    C = np.sqrt(9.81 * sg.kappa) / sg.kappa
    # tau_w of the capped rates (rate N / C * w * rho_w g, projected on the wind) = 0.4 rho_a u*^2.
    tw = np.sum(capped[0, 0] / C[:, None] * sg.weights * 1025.0 * 9.81 * np.cos(sg.phi)[None, :])
    assert tw == pytest.approx(0.4 * 1.225 * us * us, rel=1e-10)


def test_whitecapping_known_answers(oracle):
    """test/sources/source_terms.jl:236-261: uniform N = 1 on kappa = [0.5, 1, 2],
    phi = [0]: every tendency negative, growing in magnitude with kappa, the
    ratio of two bins' tendencies equal to the ratio of their cell averages of
    kappa^2; a saturation threshold above m0 gives no tendency."""
    g = grids.RectilinearGrid(size=(1, 1, 1), x=(0, 1), y=(0, 1), z=(0, 1))
    sg = grids.PolarWaveVectorGrid(kappa=[0.5, 1.0, 2.0], phi=[0.0])
    wl = workloads.Workload("wc", g, sg, None, np.ones((1, 1, 3, 1), order="F"), 1.0, 1,
                            dict(horizontal_advection=None, timestepper="ForwardEuler"))
    wc = ra.WhitecappingDissipation(rate=0.2, saturation_threshold=0.1, saturation_power=1.0, wavenumber_power=2.0,
                                    reference_wavenumber=1.0)
    G = oracle.OracleSourceModel(wl, src_struct(oracle, wc), transport=oracle.MODE_NONE).tendency()[0, 0, :, 0]
    kf, pf = sg.kappa_faces, sg.phi_faces
    w = (kf[1:] ** 2 - kf[:-1] ** 2) / 2 * (pf[1] - pf[0])
    avg_k2 = (pf[1] - pf[0]) * (kf[1:] ** 4 - kf[:-1] ** 4) / 4 / w
    assert np.all(G < 0)
    assert G[1] < G[0] and G[2] < G[1]
    assert G[2] / G[1] == pytest.approx(avg_k2[2] / avg_k2[1], rel=1e-12)
    # closed form of the first bin: rate * (m0 / threshold - 1) * <k^2> * N
    m0 = float(np.sum(w))
    assert G[0] == pytest.approx(-0.2 * (m0 / 0.1 - 1) * avg_k2[0], rel=1e-12)
    weak = ra.WhitecappingDissipation(rate=0.2, saturation_threshold=100.0)
    G = oracle.OracleSourceModel(wl, src_struct(oracle, weak), transport=oracle.MODE_NONE).tendency()
    assert np.max(np.abs(G)) < 1e-14


def test_storm_track_known_answers(oracle):
    """test/forcing/winds.jl:16-21: a three-point track, clamped at both ends."""
    times, centers = [0.0, 1.0, 3.0], [(0.0, 0.0), (2.0, 1.0), (4.0, -1.0)]
    assert oracle.track_at(times, centers, -1.0) == (0.0, 0.0)
    assert oracle.track_at(times, centers, 0.5) == (1.0, 0.5)
    assert oracle.track_at(times, centers, 2.0) == (3.0, 0.0)
    assert oracle.track_at(times, centers, 4.0) == (4.0, -1.0)
    tr = ra.LinearStormTrack(times, centers)
    for t in (-1.0, 0.0, 0.5, 1.0, 2.0, 3.0, 4.0):
        assert tr(t) == oracle.track_at(times, centers, t)
    with pytest.raises(ValueError):
        ra.LinearStormTrack([0.0, 0.0], [(0.0, 0.0), (1.0, 1.0)])
    with pytest.raises(ValueError):
        ra.LinearStormTrack([0.0], [(0.0, 0.0), (1.0, 1.0)])
    with pytest.raises(ValueError):
        ra.LinearStormTrack([0.0], [(0.0, 0.0, 0.0)])
