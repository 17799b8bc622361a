"""A spatially varying depth (SpectralWaveModel(...; depth=f(x, y)), the
reference's is_spatially_varying_depth, Models/transport.jl:102-139, with the
Q depth of a prescribed current per column, q_transform.jl:124-125) through
the C ABI against the oracle."""
import math
import sys

import numpy as np
import pytest

import ripple_amd as ra
from ripple_amd import grids, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def bathymetry(x, y):
    """A shelf: 0.3 deep at x = 0 rising to 3 at x = 80, with a bump in y."""
    return 0.3 + 2.7 * (x / 80.0) + 0.2 * np.sin(2 * np.pi * y / 80.0) ** 2


def moving_packet(wl):
    nx, ny = wl.N0.shape[:2]
    x = np.arange(nx)[:, None, None, None]
    y = np.arange(ny)[None, :, None, None]
    return np.asfortranarray(wl.N0 * (1.2 + np.sin(2 * np.pi * x / nx) * np.cos(4 * np.pi * y / ny)))


def build(wl, float_type, kernel, depth, transport):
    g, sg = wl.grid, wl.spectral_grid
    grid = grids.RectilinearGrid(float_type=float_type, size=(g.Nx, g.Ny, g.Nz), x=g.x_faces, y=g.y_faces,
                                 z=g.z_faces)
    sgrid = grids.PolarWaveVectorGrid(float_type, kappa=sg.kappa, phi=sg.phi, kappa_faces=sg.kappa_faces,
                                      phi_faces=sg.phi_faces)
    kw = dict(timestepper="RK3", kernel=kernel, depth=depth)
    if transport == "refraction":
        kw["velocities"] = wl.velocities
    m = ra.SpectralWaveModel(grid, sgrid, **kw)
    ra.set_(m, N=wl.N0)
    return m


def field_of(wl):
    X, Y = np.meshgrid(ra.xnodes(wl.grid), ra.ynodes(wl.grid), indexing="ij")
    return bathymetry(X, Y)


@pytest.mark.parametrize("transport", ["intrinsic", "refraction"])
def test_fp64_against_oracle(oracle, transport):
    say(f"depth field fp64 {transport}")
    wl = workloads.vortex_refraction(nsteps=3)
    wl.N0 = moving_packet(wl)
    d = field_of(wl)
    mode = oracle.MODE_INTRINSIC if transport == "intrinsic" else oracle.MODE_REFRACTION
    if transport == "intrinsic":
        wl.velocities = None
    om = oracle.OracleModel(wl, depth=d, transport=mode, timestepper="RK3")
    m = build(wl, np.float64, "auto", bathymetry, transport)
    if transport == "refraction":
        # the Doppler velocity from each column's Q weights at its depth
        Ux, Uy = m.doppler_velocity()
        ox, oy = om.doppler_velocity()
        say(f"U rel-L2 {rel_l2(Ux, ox):.3e}")
        assert rel_l2(Ux, ox) < 1e-13 and rel_l2(Uy, oy) < 1e-13
    G, Go = m.tendency(), om.tendency()
    err_g = rel_l2(G, Go)
    ra.time_step_(m, wl.dt, 3)
    om.step(wl.dt, 3)
    err = rel_l2(m.action, om.action)
    say(f"tendency rel-L2 {err_g:.3e}, 3 RK3 steps {err:.3e}")
    if transport == "intrinsic":
        assert np.array_equal(G, Go) and np.array_equal(m.action, om.action)  # same operations throughout
    else:
        assert err_g < 1e-12 and err < 1e-12
    # the depth matters: a constant depth at the field's mean gives another answer
    flat = oracle.OracleModel(wl, depth=float(d.mean()), transport=mode, timestepper="RK3")
    flat.step(wl.dt, 3)
    assert rel_l2(flat.action, om.action) > 1e-6


@pytest.mark.parametrize("transport", ["intrinsic", "refraction"])
def test_phi_walk_kernels(oracle, transport):
    say(f"depth field phi-walk kernels {transport}")
    wl = workloads.vortex_refraction(nsteps=3)
    wl.N0 = moving_packet(wl)
    if transport == "intrinsic":
        wl.velocities = None
    mode = oracle.MODE_INTRINSIC if transport == "intrinsic" else oracle.MODE_REFRACTION
    om = oracle.OracleModel(wl, depth=field_of(wl), transport=mode, timestepper="RK3")
    om.step(wl.dt, 3)
    m64 = build(wl, np.float64, "fast", bathymetry, transport)
    ra.time_step_(m64, wl.dt, 3)
    e64 = rel_l2(m64.action, om.action)
    # FP32 against the FP32 oracle with the bounds of the constant-depth
    # tests (test_gpu_parity.py: intrinsic 5e-6, refraction 5e-5)
    m32 = build(wl, np.float32, "auto", bathymetry, transport)
    ra.time_step_(m32, wl.dt, 3)
    o32 = oracle.OracleModel(wl, dtype=np.float32, depth=field_of(wl), transport=mode, timestepper="RK3")
    o32.step(wl.dt, 3)
    e32 = rel_l2(m32.action, o32.action)
    say(f"FP64 fast vs FP64 oracle {e64:.3e}, FP32 vs FP32 oracle {e32:.3e}, "
        f"FP32 vs FP64 oracle {rel_l2(m32.action, om.action):.3e}")
    assert e64 < 1e-12
    assert e32 < (5e-6 if transport == "intrinsic" else 5e-5)


def test_uniform_field_is_the_constant_depth():
    say("depth field equal to a constant")
    wl = workloads.vortex_refraction(nsteps=2)
    wl.N0 = moving_packet(wl)
    wl.velocities = None
    a = build(wl, np.float64, "auto", lambda x, y: 1.5 + 0 * x, "intrinsic")
    b = build(wl, np.float64, "auto", 1.5, "intrinsic")
    ra.time_step_(a, wl.dt, 2)
    ra.time_step_(b, wl.dt, 2)
    assert np.array_equal(a.action, b.action)


def test_depth_errors():
    say("depth field errors")
    wl = workloads.vortex_refraction(nsteps=1)
    with pytest.raises(ra.RippleArgumentError):
        build(wl, np.float64, "auto", np.ones((64, 64)), "intrinsic")  # raw arrays are not accepted
    with pytest.raises(ra.RippleArgumentError):
        build(wl, np.float64, "auto", lambda x, y: x - 40.0, "intrinsic")  # values must be positive
    with pytest.raises(ra.RippleArgumentError):
        build(wl, np.float64, "auto", ra.HorizontalField(np.ones((3, 3))), "intrinsic")
    m = build(wl, np.float64, "auto", ra.HorizontalField(np.full((64, 64), 2.0)), "intrinsic")
    assert m.depth_field.shape == (64, 64)
    a = build(wl, np.float32, "auto", bathymetry, "refraction")
    b = build(wl, np.float32, "auto", None, "refraction")
    with pytest.raises(ra.RippleError):
        ra.Ensemble([a, b])


def shelf(x, y):
    """4 m at x = 0 rising to 12 m at x = 80 (the packet's box), with a bump."""
    return 4.0 + 8.0 * (x / 80.0) + 1.0 * np.cos(2 * np.pi * y / 80.0) ** 2


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_pseudomomentum_with_depth_field(oracle, dtype):
    """pseudomomentum.jl:72-87 with a depth per column: the overlap tables of
    every column at its depth, the per-stage U and dU/dkappa, two RK3 steps
    and the cell averages (q_depth_at, pseudomomentum.jl:375-400) against
    the oracle's per-column tables (device FP64 exp: rounding-level)."""
    say(f"pseudomomentum with a depth field {np.dtype(dtype).name}")
    wl = workloads.pseudomomentum_packet(float_type=dtype, nx=32, ny=32, nz=8, nkappa=6, nphi=16)
    g, sg = wl.grid, wl.spectral_grid
    X, Y = np.meshgrid(ra.xnodes(g), ra.ynodes(g), indexing="ij")
    d = shelf(X, Y)
    m = ra.SpectralWaveModel(g, sg, depth=shelf, **wl.model_kwargs)
    ra.set_(m, N=wl.N0)
    ra.update_coupling_(m)
    O, dO = oracle.pm_overlap_columns(g.z_faces, sg.kappa, d)
    t = oracle.geometry(sg.kappa, sg.kappa_faces, sg.phi, sg.phi_faces)
    want = oracle.pm_doppler(wl.N0, O, dO, t["mu_x"], t["mu_y"])
    got = list(m.doppler_velocity()) + list(m.doppler_velocity_kappa_derivative())
    tol_u = 1e-13 if dtype == np.float64 else 2e-6
    for name, a, b in zip(("Ux", "Uy", "dUx", "dUy"), got, want):
        say(f"{name} rel-L2 {rel_l2(a, b):.3e}")
        assert rel_l2(a, b) < tol_u, name
    # the depth matters: one table at the mean depth gives another velocity
    Om, dOm = oracle.pm_overlap(g.z_faces, sg.kappa, float(d.mean()))
    assert rel_l2(got[0], oracle.pm_doppler(wl.N0, Om, dOm, t["mu_x"], t["mu_y"])[0]) > 1e-4
    P = ra.compute_pseudomomentum_cell_averages(m)
    Pw = oracle.pm_cells(wl.N0, oracle.q_weights_columns(g.z_faces, sg.kappa, d), g.z_faces, t["mu_x"], t["mu_y"])
    say(f"cell averages rel-L2 {rel_l2(P[0], Pw[0]):.3e}")
    assert rel_l2(P[0], Pw[0]) < tol_u and rel_l2(P[1], Pw[1]) < tol_u
    om = oracle.OracleModel(wl, dtype=dtype, depth=d, pseudomomentum=(g.z_faces, g.Lz))
    ra.time_step_(m, wl.dt, 2)
    om.step(wl.dt, 2)
    err = rel_l2(m.action, om.action)
    say(f"2 RK3 steps rel-L2 {err:.3e}")
    assert err < (1e-12 if dtype == np.float64 else 5e-5)
