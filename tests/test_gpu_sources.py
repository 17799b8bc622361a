"""Device source terms against the FP64 oracle (wind input with stress cap,
mean-spectrum and local-saturation whitecapping, uniform and Holland winds)."""
import math
import sys

import numpy as np
import pytest

import ripple_amd as ra
from ripple_amd import grids, workloads
from ripple_amd.sources import to_struct

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def device(wl, float_type, kernel="auto"):
    g, sg = wl.grid, wl.spectral_grid
    grid = grids.RectilinearGrid(float_type=float_type, size=(g.Nx, g.Ny, g.Nz), x=g.x_faces, y=g.y_faces,
                                 z=g.z_faces)
    if isinstance(sg, grids.FrequencyDirectionGrid):
        sgrid = grids.FrequencyDirectionGrid(float_type, frequency=sg.frequency, phi=sg.phi,
                                             frequency_faces=sg.frequency_faces, phi_faces=sg.phi_faces)
    else:
        sgrid = grids.PolarWaveVectorGrid(float_type, kappa=sg.kappa, phi=sg.phi, kappa_faces=sg.kappa_faces,
                                          phi_faces=sg.phi_faces)
    m = ra.SpectralWaveModel(grid, sgrid, velocities=wl.velocities, kernel=kernel, **wl.model_kwargs)
    ra.set_(m, N=wl.N0)
    return m


def oracle_model(oracle, wl):
    src = oracle.sources_from(to_struct(wl.model_kwargs["sources"]))
    return oracle.OracleSourceModel(wl, src, depth=wl.model_kwargs.get("depth"))


def test_holland_wind_columns(oracle):
    say("holland wind columns")
    wl = workloads.translating_hurricane(float_type=np.float64, nx=48, ny=40, nz=4, nkappa=6, nphi=8)
    m = device(wl, np.float64)
    src = oracle.sources_from(to_struct(wl.model_kwargs["sources"]))
    for t in (0.0, 3600.0 * 7):
        m.clock.time = t
        ra._lib.check(m._ctx.lib.rpl_set_clock(m._ctx.ptr, t, 0))
        U, th, us = ra.wind_state(m)
        Uo, tho, uso = oracle.wind_columns(src, grids.xnodes(wl.grid), grids.ynodes(wl.grid), t)
        np.testing.assert_allclose(U, Uo, rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(us, uso, rtol=1e-13, atol=1e-15)
        assert np.max(np.abs(np.angle(np.exp(1j * (th - tho))))) < 1e-12


def test_holland_wind_on_a_multi_point_track(oracle):
    """LinearStormTrack of four points (winds.jl:4-41) on the device against
    the oracle's track, at knots, inside segments and past both ends."""
    say("holland wind, 4-point track")
    wl = workloads.translating_hurricane(float_type=np.float64, nx=48, ny=40, nz=4, nkappa=6, nphi=8)
    w = wl.model_kwargs["sources"].wind_input.wind
    times = [3600.0, 7200.0, 14400.0, 30000.0]
    centers = [(150e3, -50e3), (250e3, 20e3), (300e3, 120e3), (500e3, 0.0)]
    w.track = ra.LinearStormTrack(times, centers)
    m = device(wl, np.float64)
    src = oracle.set_track(oracle.sources_from(to_struct(wl.model_kwargs["sources"])), times, centers)
    for t in (0.0, 3600.0, 5000.0, 7200.0, 10000.0, 29999.0, 40000.0):
        ra._lib.check(m._ctx.lib.rpl_set_clock(m._ctx.ptr, t, 0))
        U, th, us = ra.wind_state(m)
        Uo, tho, uso = oracle.wind_columns(src, grids.xnodes(wl.grid), grids.ynodes(wl.grid), t)
        np.testing.assert_allclose(U, Uo, rtol=1e-13, atol=1e-13)
        assert np.max(np.abs(np.angle(np.exp(1j * (th - tho))))) < 1e-12


@pytest.mark.parametrize("case", ["fetch", "hurricane"])
def test_source_tendency_and_steps_fp64(oracle, case):
    say(f"sources fp64 {case}")
    if case == "fetch":
        wl = workloads.fetch_limited_growth(float_type=np.float64, nx=24, ny=20, nz=4, nfreq=12, nphi=12, dt=10.0)
    else:
        wl = workloads.translating_hurricane(float_type=np.float64, nx=40, ny=36, nz=4, nkappa=10, nphi=12, dt=60.0)
        # Put the storm inside this small box.
        s = wl.model_kwargs["sources"].wind_input.wind
        s.rmax, s.radius = 3e3, 30e3
        s.track.centers = [(200e3, 0.0), (200e3 + 5.0 * s.track.times[1], 0.0)]
    m = device(wl, np.float64)
    om = oracle_model(oracle, wl)
    G, Go = m.tendency(), om.tendency()
    assert rel_l2(G, Go) <= 1e-12, rel_l2(G, Go)
    for _ in range(3):
        ra.time_step_(m, wl.dt)
        om.step(wl.dt)
    err = rel_l2(m.action, om.action)
    say(f"sources fp64 {case}: rel-L2 after 3 steps {err:.2e}")
    assert err <= 1e-12
    assert m.clock.time == pytest.approx(om.time)


@pytest.mark.parametrize("case", ["fetch", "hurricane", "hurricane_powers"])
def test_source_steps_fp32_fast(oracle, case):
    """FP32 stage kernels (tabulated wind-input factors, sigma k^3 and sigma^q
    tables, squares for exponent 2) against the FP64 oracle; the last case
    takes the general-power branches (wind-input cosine power 3, saturation
    exponent 1.5, sigma power 1.3)."""
    say(f"sources fp32 {case}")
    if case == "fetch":
        wl = workloads.fetch_limited_growth(float_type=np.float64, nx=24, ny=20, nz=4, nfreq=12, nphi=12, dt=10.0)
    else:
        wl = workloads.translating_hurricane(float_type=np.float64, nx=40, ny=36, nz=4, nkappa=10, nphi=12, dt=60.0)
    if case == "hurricane_powers":
        win = wl.model_kwargs["sources"].wind_input
        wl.model_kwargs["sources"] = ra.PrecomputedSources(
            wind_input=ra.PressureCorrelationInput(drag=win.drag, wind=win.wind, p_in=3.0),
            dissipation=ra.LocalSaturationDissipation(B_r=1.05e-2, p=1.5, sigma_power=1.3))
    m = device(wl, np.float32, kernel="fast")
    om = oracle_model(oracle, wl)
    for _ in range(5):
        ra.time_step_(m, wl.dt)
        om.step(wl.dt)
    err = rel_l2(m.action, om.action)
    say(f"sources fp32 {case}: rel-L2 vs fp64 after 5 steps {err:.2e}")
    assert err <= 2e-3


@pytest.mark.parametrize("bundle", [True, False])
def test_whitecapping_fp64_parity(oracle, bundle):
    """WhitecappingDissipation (Sources.jl:583-598) on the device against the
    oracle, inside the bundle with the wind input and on its own."""
    say(f"whitecapping fp64 bundle={bundle}")
    wl = workloads.fetch_limited_growth(float_type=np.float64, nx=20, ny=18, nz=4, nfreq=10, nphi=12, dt=10.0)
    wc = ra.WhitecappingDissipation(rate=2e-4, saturation_threshold=1e-5, saturation_power=1.5,
                                    wavenumber_power=2.0, reference_wavenumber=0.5)
    if bundle:
        wl.model_kwargs["sources"] = ra.PrecomputedSources(
            wind_input=wl.model_kwargs["sources"].wind_input, dissipation=wc)
    else:
        wl.model_kwargs["sources"] = wc
    m = device(wl, np.float64)
    om = oracle_model(oracle, wl)
    G, Go = m.tendency(), om.tendency()
    e0 = rel_l2(G, Go)
    for _ in range(3):
        ra.time_step_(m, wl.dt)
        om.step(wl.dt)
    e = rel_l2(m.action, om.action)
    say(f"whitecapping bundle={bundle}: tendency rel-L2 {e0:.2e}, after 3 steps {e:.2e}")
    assert e0 <= 1e-12 and e <= 1e-12
    m32 = device(wl, np.float32)
    ra.time_step_(m32, wl.dt, 3)
    assert rel_l2(m32.action, om.action) <= 2e-3


def test_fetch_limited_growth_full_size_properties():
    """Config 3 at its full size in FP32: waves grow under a 10 m/s wind, stay
    finite and non-negative, and the spectrum's energy moves toward the wind
    direction."""
    say("fetch-limited full size fp32")
    wl = workloads.fetch_limited_growth()
    m = device(wl, np.float32)
    hs0 = float(np.mean(ra.significant_wave_height(m)))
    ra.time_step_(m, wl.dt, 60)
    N = m.action
    assert np.all(np.isfinite(N)) and np.all(N >= 0)
    hs = float(np.mean(ra.significant_wave_height(m)))
    say(f"fetch-limited: mean Hs {hs0:.4f} -> {hs:.4f} after 10 min")
    assert hs > hs0
    assert abs(float(np.mean(ra.mean_direction(m)))) < 0.05


def test_fused_stress_cap_fp32(oracle):
    """The FP32 fast kernels take the stress cap of RK3 stages 2 and 3 from
    the wave-supported stress their previous stage summed for its output
    (stage 1 runs the column-state kernel).  On a strong sea under 25 m/s,
    where the cap scales the growth down, the FP32 run matches the FP32
    reference-order kernel (column-state kernel every stage) to rounding and
    the FP64 oracle to the FP32 tolerance."""
    say("fused stress cap fp32")
    g = grids.RectilinearGrid(size=(70, 9, 2), x=(0, 70e3), y=(0, 9e3), z=(-50.0, 0.0))
    sg = grids.PolarWaveVectorGrid(kappa=np.geomspace(0.02, 0.6, 10), phi=np.linspace(0, 2 * math.pi, 13)[:12])
    inp = ra.PressureCorrelationInput(drag=ra.BulkWindDrag("linear"), wind=25.0, direction=0.4)
    src = ra.PrecomputedSources(wind_input=inp, dissipation=ra.LocalSaturationDissipation(B_r=1.05e-2))
    rng = np.random.default_rng(11)
    N0 = np.asfortranarray(40.0 * (1 + 0.5 * rng.random((70, 9, 10, 12))))
    wl = workloads.Workload("cap", g, sg, None, N0, 20.0, 3, dict(timestepper="RK3", sources=src))
    # the cap is active: the capped growth is well below the raw growth
    om = oracle_model(oracle, wl)
    raw = oracle.OracleSourceModel(wl, oracle.sources_from(to_struct(inp)), transport=oracle.MODE_NONE).tendency()
    capped = oracle.OracleSourceModel(
        wl, oracle.sources_from(to_struct(ra.PrecomputedSources(wind_input=inp))),
        transport=oracle.MODE_NONE).tendency()
    assert capped.sum() < 0.7 * raw.sum()
    fast = device(wl, np.float32, kernel="fast")
    ref = device(wl, np.float32, kernel="reference")
    for _ in range(3):
        ra.time_step_(fast, wl.dt)
        ra.time_step_(ref, wl.dt)
        om.step(wl.dt)
    d_ref = rel_l2(fast.action, ref.action)
    d_orc = rel_l2(fast.action, om.action)
    say(f"fused cap fp32: vs fp32 reference kernel {d_ref:.2e}, vs fp64 oracle {d_orc:.2e}")
    assert d_ref <= 2e-5 and d_orc <= 1e-4


def test_sources_fp64_fast_kernel(oracle):
    """Intrinsic transport with the source bundle in FP64 on the phi-loop
    kernel: within 1e-12 of the oracle after three RK3 steps."""
    say("sources fp64 fast kernel")
    wl = workloads.translating_hurricane(float_type=np.float64, nx=40, ny=36, nz=4, nkappa=10, nphi=12, dt=60.0)
    s = wl.model_kwargs["sources"].wind_input.wind
    s.rmax, s.radius = 3e3, 30e3
    s.track.centers = [(200e3, 0.0), (200e3 + 5.0 * s.track.times[1], 0.0)]
    m = device(wl, np.float64, kernel="fast")
    om = oracle_model(oracle, wl)
    assert rel_l2(m.tendency(), om.tendency()) <= 1e-12
    for _ in range(3):
        ra.time_step_(m, wl.dt)
        om.step(wl.dt)
    err = rel_l2(m.action, om.action)
    say(f"sources fp64 fast: rel-L2 after 3 steps {err:.2e}")
    assert err <= 1e-12
