"""Device parity of SpatialAveraging (Models/propagation_smoothing.jl), the
propagation smoothing applied after every time step, through the C ABI,
against the oracle: FP64 bit-identical, FP32 within the tolerance written in
each test."""
import dataclasses
import math
import sys

import numpy as np
import pytest

import ripple_amd as ra
from ripple_amd import workloads, grids

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def model(wl, float_type, transport, alphas=(0.5, 0.5), kernel="auto"):
    g, sg = wl.grid, wl.spectral_grid
    grid = grids.RectilinearGrid(float_type=float_type, size=(g.Nx, g.Ny, g.Nz), x=g.x_faces, y=g.y_faces,
                                 z=g.z_faces, topology=g.topology)
    sgrid = grids.PolarWaveVectorGrid(float_type, kappa=sg.kappa, phi=sg.phi, kappa_faces=sg.kappa_faces,
                                      phi_faces=sg.phi_faces)
    kw = dict(timestepper="RK3", kernel=kernel, propagation_smoothing=ra.SpatialAveraging(*alphas))
    if transport == "none":
        kw["horizontal_advection"] = None
    elif transport == "refraction":
        kw["velocities"] = wl.velocities
    m = ra.SpectralWaveModel(grid, sgrid, **kw)
    ra.set_(m, N=wl.N0)
    return m


def packet(**kw):
    return workloads.pseudomomentum_packet(**kw)


@pytest.mark.parametrize("transport", ["none", "intrinsic"])
def test_fp64_bitwise(oracle, transport):
    say(f"spatial averaging fp64 transport={transport}")
    wl = packet(nx=64, ny=48, nkappa=8, nphi=24)
    dt = 5.0 if transport == "none" else wl.dt  # no transport: any step; the averaging moves cells
    mode = oracle.MODE_NONE if transport == "none" else oracle.MODE_INTRINSIC
    m = model(wl, np.float64, transport)
    om = oracle.OracleModel(wl, transport=mode, timestepper="RK3", smoothing=(0.5, 0.5))
    ra.time_step_(m, dt, 3)
    om.step(dt, 3)
    say(f"rel-L2 {rel_l2(m.action, om.action):.3e}")
    assert np.array_equal(m.action, om.action)
    assert m.clock.iteration == 3 and m.clock.time == pytest.approx(3 * dt)
    if transport == "none":
        # The averaging alone moved the field, and kept every bin's total.
        assert rel_l2(m.action, wl.N0) > 1e-3
        np.testing.assert_allclose(m.action.sum(axis=(0, 1)), wl.N0.sum(axis=(0, 1)), rtol=1e-12)


def test_fp64_refraction_bitwise(oracle):
    say("spatial averaging fp64 with refraction (vortex)")
    wl = workloads.vortex_refraction(nsteps=2)
    m = model(wl, np.float64, "refraction", alphas=(1.0, 0.75))
    om = oracle.OracleModel(wl, timestepper="RK3", smoothing=(1.0, 0.75))
    ra.time_step_(m, wl.dt, 2)
    om.step(wl.dt, 2)
    assert np.array_equal(m.action, om.action)


@pytest.mark.parametrize("kernel", ["auto", "reference"])
def test_fp32_against_fp64_oracle(oracle, kernel):
    say(f"spatial averaging fp32 kernel={kernel}, 5 steps")
    wl = packet(nx=64, ny=64, nkappa=12, nphi=24)
    m = model(wl, np.float32, "intrinsic", kernel=kernel)
    om = oracle.OracleModel(wl, transport=oracle.MODE_INTRINSIC, timestepper="RK3", smoothing=(0.5, 0.5))
    ra.time_step_(m, wl.dt, 5)
    om.step(wl.dt, 5)
    err = rel_l2(m.action, om.action)
    say(f"FP32 vs FP64 oracle rel-L2 {err:.3e}")
    assert err < 2e-6


def test_large_shape_conserves_each_bin():
    """256 x 256 x 32 x 36 FP32 without transport: the averaging keeps every
    bin's total (periodic grid) and leaves a uniform field uniform."""
    say("spatial averaging fp32 256^2 x 32 x 36 conservation")
    wl = packet(float_type=np.float32, nx=256, ny=256, nz=4, nkappa=32, nphi=36)
    m = model(wl, np.float32, "none")
    before = m.action.astype(np.float64).sum(axis=(0, 1))
    ra.time_step_(m, 10.0, 2)
    after = m.action.astype(np.float64).sum(axis=(0, 1))
    np.testing.assert_allclose(after, before, rtol=2e-5)
    ra.set_(m, N=np.full(wl.N0.shape, 0.25))
    ra.time_step_(m, 10.0, 1)
    np.testing.assert_allclose(m.action, 0.25, rtol=1e-6)


def test_smoothing_error_paths():
    say("spatial averaging errors")
    wl = packet(nx=16, ny=16, nkappa=4, nphi=8)
    with pytest.raises(ra.RippleArgumentError):
        model(wl, np.float64, "intrinsic", alphas=(-1.0, 0.5))
    with pytest.raises(TypeError):
        g, sg = wl.grid, wl.spectral_grid
        ra.SpectralWaveModel(g, sg, propagation_smoothing="Tolman")
    a = model(wl, np.float32, "intrinsic")
    b = model(wl, np.float32, "intrinsic")
    with pytest.raises(ra.RippleError):
        ra.Ensemble([a, b])


def bounded(wl, tx, ty):
    """The workload on a grid with x and y topologies tx, ty."""
    g = wl.grid
    grid = grids.RectilinearGrid(float_type=g.float_type, size=(g.Nx, g.Ny, g.Nz), x=g.x_faces, y=g.y_faces,
                                 z=g.z_faces, topology=(tx, ty, grids.BOUNDED))
    return dataclasses.replace(wl, grid=grid)


@pytest.mark.parametrize("tx,ty", [(grids.BOUNDED, grids.BOUNDED), (grids.PERIODIC, grids.BOUNDED)])
def test_fp64_bounded_refraction_bitwise(oracle, tx, ty):
    """Refraction wraps x and y on any topology (wave_current_refraction.jl:
    97-111); the averaging clamps its corners on a bounded axis
    (propagation_smoothing.jl:97-100)."""
    say(f"spatial averaging fp64 refraction on a {tx} x {ty} grid")
    wl = bounded(workloads.vortex_refraction(nsteps=2), tx, ty)
    m = model(wl, np.float64, "refraction", alphas=(1.0, 0.75))
    om = oracle.OracleModel(wl, timestepper="RK3", smoothing=(1.0, 0.75))
    ra.time_step_(m, wl.dt, 2)
    om.step(wl.dt, 2)
    assert np.array_equal(m.action, om.action)
    # the clamp is what differs from the periodic run
    per = oracle.OracleModel(workloads.vortex_refraction(nsteps=2), timestepper="RK3", smoothing=(1.0, 0.75))
    per.step(wl.dt, 2)
    assert not np.array_equal(m.action, per.action)


def test_fp64_bounded_no_transport_bitwise(oracle):
    say("spatial averaging fp64 without transport on a Bounded x Periodic grid")
    wl = bounded(packet(nx=64, ny=48, nkappa=8, nphi=24), grids.BOUNDED, grids.PERIODIC)
    m = model(wl, np.float64, "none")
    om = oracle.OracleModel(wl, transport=oracle.MODE_NONE, timestepper="RK3", smoothing=(0.5, 0.5))
    ra.time_step_(m, 5.0, 3)
    om.step(5.0, 3)
    assert np.array_equal(m.action, om.action)
    assert rel_l2(m.action, wl.N0) > 1e-3


@pytest.mark.parametrize("kernel", ["auto", "reference"])
@pytest.mark.parametrize("tx,ty", [(grids.BOUNDED, grids.BOUNDED), (grids.BOUNDED, grids.PERIODIC),
                                   (grids.PERIODIC, grids.BOUNDED)])
def test_fp32_bounded_against_fp64_oracle(oracle, kernel, tx, ty):
    say(f"spatial averaging fp32 kernel={kernel} on a {tx} x {ty} grid, 5 steps")
    wl = bounded(packet(nx=64, ny=64, nkappa=12, nphi=24), tx, ty)
    m = model(wl, np.float32, "none", kernel=kernel, alphas=(1.5, 0.75))
    om = oracle.OracleModel(wl, transport=oracle.MODE_NONE, timestepper="RK3", smoothing=(1.5, 0.75))
    ra.time_step_(m, 20.0, 5)
    om.step(20.0, 5)
    err = rel_l2(m.action, om.action)
    say(f"FP32 vs FP64 oracle rel-L2 {err:.3e}")
    assert err < 2e-6
    assert rel_l2(m.action, wl.N0) > 1e-3


def test_topology_errors():
    say("grid topology errors")
    wl = bounded(packet(nx=16, ny=16, nkappa=4, nphi=8), grids.BOUNDED, grids.PERIODIC)
    with pytest.raises(NotImplementedError):
        model(wl, np.float64, "intrinsic")  # tendencies.jl:45-50: Oceananigans' per-bin path
    m = model(packet(nx=16, ny=16, nkappa=4, nphi=8), np.float64, "intrinsic")
    lib = m._ctx.lib
    with pytest.raises(ra.RippleError):
        ra._lib.check(lib.rpl_set_topology(m._ctx.ptr, 0, 1))
    with pytest.raises(ra.RippleArgumentError):
        ra._lib.check(lib.rpl_set_topology(m._ctx.ptr, 2, 1))
    ra._lib.check(lib.rpl_set_topology(m._ctx.ptr, 1, 1))
