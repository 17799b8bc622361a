"""Device parity of the self-coupled pseudomomentum current
(CWCMPseudomomentumCoupling, Coupling/pseudomomentum.jl) and of the Doppler
kappa derivatives, through the C ABI, against the oracle: FP64 reference-order
contractions bit-identical, FP32 and the moment-factored kernels within the
tolerances written in each test."""
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


def packet(**kw):
    return workloads.pseudomomentum_packet(**kw)


def device_model(wl, float_type, kernel="auto", **extra):
    g = wl.grid
    grid = grids.RectilinearGrid(float_type=float_type, size=(g.Nx, g.Ny, g.Nz), x=g.x_faces, y=g.y_faces,
                                 z=g.z_faces)
    sg = wl.spectral_grid
    sgrid = grids.PolarWaveVectorGrid(float_type, kappa=sg.kappa, phi=sg.phi, kappa_faces=sg.kappa_faces,
                                      phi_faces=sg.phi_faces)
    kw = dict(wl.model_kwargs)
    kw.update(extra)
    m = ra.SpectralWaveModel(grid, sgrid, kernel=kernel, **kw)
    ra.set_(m, N=wl.N0)
    return m


def oracle_model(oracle, wl, dtype=np.float64):
    g = wl.grid
    return oracle.OracleModel(wl, dtype=dtype, pseudomomentum=(g.z_faces, g.Lz))


def oracle_velocity(oracle, wl, N):
    g, sg = wl.grid, wl.spectral_grid
    t = oracle.geometry(sg.kappa, sg.kappa_faces, sg.phi, sg.phi_faces)
    O, dO = oracle.pm_overlap(g.z_faces, sg.kappa, g.Lz)
    return oracle.pm_doppler(N, O, dO, t["mu_x"], t["mu_y"])


def test_coupling_velocity_fp64_bitwise(oracle):
    say("pseudomomentum U, dU/dkappa fp64 packet 64x64x16 k8 p24")
    wl = packet()
    m = device_model(wl, np.float64)
    Ux, Uy = m.doppler_velocity()
    assert not Ux.any() and not Uy.any()  # a new coupling holds zeros (current_coupling.jl:104-109)
    ra.update_coupling_(m)
    want = oracle_velocity(oracle, wl, wl.N0)
    got = list(m.doppler_velocity()) + list(m.doppler_velocity_kappa_derivative())
    for name, a, b in zip(("Ux", "Uy", "dUx", "dUy"), got, want):
        say(f"{name}: max |dev - oracle| = {np.abs(a - b).max():.3e}")
        assert np.array_equal(a, b), name


def test_coupling_velocity_linear_in_action(oracle):
    """model_api.jl:232-235: doubling the action doubles U (exactly)."""
    say("pseudomomentum linearity")
    wl = packet(nx=32, ny=32)
    for ft, kernel in ((np.float64, "auto"), (np.float32, "auto"), (np.float64, "fast")):
        m = device_model(wl, ft, kernel=kernel)
        ra.update_coupling_(m)
        U1 = m.doppler_velocity()
        ra.set_(m, N=2 * wl.N0)
        ra.update_coupling_(m)
        U2 = m.doppler_velocity()
        for a, b in zip(U1, U2):
            assert np.array_equal(2 * a, b), (ft, kernel)


def test_tendency_and_steps_fp64_bitwise(oracle):
    say("pseudomomentum tendency + 3 RK3 steps fp64 packet 64x64")
    wl = packet()
    m = device_model(wl, np.float64)
    om = oracle_model(oracle, wl)
    G = m.tendency()
    Gw = om.tendency()
    say(f"tendency rel-L2 {rel_l2(G, Gw):.3e}")
    assert np.array_equal(G, Gw)
    ra.time_step_(m, wl.dt, 3)
    om.step(wl.dt, 3)
    err = rel_l2(m.action, om.action)
    say(f"3 steps rel-L2 {err:.3e}")
    assert np.array_equal(m.action, om.action)
    # U now holds the velocity of the last stage input, as the reference's cache.
    Ux, Uy = m.doppler_velocity()
    oUx, oUy = om.doppler_velocity()
    assert np.array_equal(Ux, oUx) and np.array_equal(Uy, oUy)


@pytest.mark.parametrize("kernel", ["auto", "fast"])
def test_fp32_against_fp64_oracle(oracle, kernel):
    say(f"pseudomomentum fp32 kernel={kernel} vs fp64 oracle, 10 steps")
    wl = packet(nx=64, ny=48, nkappa=12, nphi=36)
    m = device_model(wl, np.float32, kernel=kernel)
    ra.update_coupling_(m)
    U = m.doppler_velocity()
    want = oracle_velocity(oracle, wl, wl.N0)
    for a, b in zip(U, want[:2]):
        err = rel_l2(a, b)
        say(f"U rel-L2 {err:.3e}")
        assert err < 2e-6
    om = oracle_model(oracle, wl)
    ra.time_step_(m, wl.dt, 10)
    om.step(wl.dt, 10)
    err = rel_l2(m.action, om.action)
    say(f"FP32 10 steps vs FP64 oracle rel-L2 {err:.3e}")
    assert err < 2e-5


def test_fp64_moment_kernel(oracle):
    """kernel="fast" on FP64 sums through the column moments: rounding-level
    agreement with the reference order."""
    say("pseudomomentum fp64 moment-factored contraction")
    wl = packet(nx=40, ny=24, nkappa=16, nphi=30)
    m = device_model(wl, np.float64, kernel="fast")
    ra.update_coupling_(m)
    want = oracle_velocity(oracle, wl, wl.N0)
    got = list(m.doppler_velocity()) + list(m.doppler_velocity_kappa_derivative())
    for a, b in zip(got, want):
        assert rel_l2(a, b) < 1e-13
    om = oracle_model(oracle, wl)
    ra.time_step_(m, wl.dt, 2)
    om.step(wl.dt, 2)
    assert rel_l2(m.action, om.action) < 1e-12


@pytest.mark.parametrize("average", [True, False])
def test_pseudomomentum_cells(oracle, average):
    say(f"pseudomomentum cells average={average}")
    wl = packet(nx=24, ny=20, nz=10)
    g, sg = wl.grid, wl.spectral_grid
    t = oracle.geometry(sg.kappa, sg.kappa_faces, sg.phi, sg.phi_faces)
    W = oracle.q_weights(g.z_faces, sg.kappa, g.Lz)
    want = oracle.pm_cells(wl.N0, W, g.z_faces, t["mu_x"], t["mu_y"], cell_average=average)
    m = device_model(wl, np.float64)
    fn = ra.compute_pseudomomentum_cell_averages if average else ra.compute_pseudomomentum_cell_integrals
    got = fn(m)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    m32 = device_model(wl, np.float32)
    for a, b in zip(fn(m32), want):
        assert rel_l2(a, b) < 2e-6
    # An explicit, different Q grid and depth.
    qz = np.array([-6.0, -3.0, -1.0, -0.25, 0.0])
    qg = grids.RectilinearGrid(size=(g.Nx, g.Ny, 4), x=g.x_faces, y=g.y_faces, z=qz)
    W2 = oracle.q_weights(qz, sg.kappa, 6.0)
    want2 = oracle.pm_cells(wl.N0, W2, qz, t["mu_x"], t["mu_y"], cell_average=average)
    for a, b in zip(fn(m, depth=6.0, q_grid=qg), want2):
        assert np.array_equal(a, b)


def test_pseudomomentum_tendency_fields(oracle):
    """pseudomomentum.jl:219-283: cell averages of (new - old) / dt, and the
    CWCM momentum forcing (coefficient -1), against the oracle's contraction
    of the same difference (FP64 bitwise)."""
    say("pseudomomentum tendency fields")
    wl = packet(nx=24, ny=20, nz=10)
    g, sg = wl.grid, wl.spectral_grid
    m = device_model(wl, np.float64)
    old = m.action.copy()
    ra.time_step_(m, wl.dt, 1)
    new = m.action
    dt = wl.dt
    t = oracle.geometry(sg.kappa, sg.kappa_faces, sg.phi, sg.phi_faces)
    W = oracle.q_weights(g.z_faces, sg.kappa, g.Lz)
    want = oracle.pm_cells((new - old) / dt, W, g.z_faces, t["mu_x"], t["mu_y"], cell_average=True)
    got = ra.compute_pseudomomentum_tendency_cell_averages(m, old, dt)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
        assert np.abs(b).max() > 0
    forcing = ra.cwcm_momentum_tendency_fields(m, old, dt)
    for a, b in zip(forcing, want):
        assert np.array_equal(a, -b)
    with pytest.raises(ra.RippleArgumentError):
        ra.compute_pseudomomentum_tendency_cell_averages(m, old, 0.0)
    # The action is untouched by the call (the stage buffers hold old and the tendency).
    assert np.array_equal(m.action, new)


def test_prescribed_current_kappa_derivative(oracle):
    """doppler_velocity.jl:25-43: dU/dkappa of a prescribed current, and the
    reference's check that a depth-uniform current has none
    (test/coupling/q_transform.jl:136-143)."""
    say("prescribed dU/dkappa")
    nx, ny, nz = 12, 10, 16
    g = grids.RectilinearGrid(size=(nx, ny, nz), x=(0, 6), y=(0, 5), z=(-1, 0))
    sg = grids.PolarWaveVectorGrid(kappa=np.geomspace(0.2, 1.5, 5), phi=np.linspace(0, 2 * math.pi, 9)[:8])
    rng = np.random.default_rng(1)
    u = np.asfortranarray(rng.normal(size=(nx, ny, nz)))
    v = np.asfortranarray(rng.normal(size=(nx, ny, nz)))
    m = ra.SpectralWaveModel(g, sg, velocities=(u, v), timestepper="RK3")
    dW = np.array([[oracle.q_cell_integral_dkappa(sg.kappa[k], g.z_faces[z], g.z_faces[z + 1], 1.0)
                    for k in range(5)] for z in range(nz)])
    want = oracle.doppler(u, v, dW)
    got = m.doppler_velocity_kappa_derivative()
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    m2 = ra.SpectralWaveModel(g, sg, velocities=(np.ones((nx, ny, nz)), 2 * np.ones((nx, ny, nz))),
                              timestepper="RK3")
    dx, dy = m2.doppler_velocity_kappa_derivative()
    assert max(np.abs(dx).max(), np.abs(dy).max()) < 1e-10
    Ux, Uy = m2.doppler_velocity()
    assert Ux[0, 0, 0] == pytest.approx(1, abs=1e-12) and Uy[1, 1, 1] == pytest.approx(2, abs=1e-12)


def test_pseudomomentum_error_paths():
    say("pseudomomentum errors")
    wl = packet(nx=16, ny=16)
    g, sg = wl.grid, wl.spectral_grid
    with pytest.raises(ValueError):
        ra.PseudomomentumVelocities(depth=2.0)
    with pytest.raises(ValueError):
        ra.PseudomomentumVelocities(q_grid=g, vertical_grid=g)
    flat = grids.RectilinearGrid(size=(16, 16), x=(0, 80), y=(0, 80))
    with pytest.raises(ValueError):  # lagrangian_velocities.jl:297-299
        ra.SpectralWaveModel(flat, sg, velocities=ra.PseudomomentumVelocities())
    # A flat grid with a finite depth takes the automatic Q grid (lagrangian_velocities.jl:196-225).
    m = ra.SpectralWaveModel(flat, sg, velocities=ra.PseudomomentumVelocities(Nz=7), depth=2.0,
                             timestepper="RK3")
    assert m.q_grid.Nz == 7 and m.q_grid.z_faces[0] == pytest.approx(-2.0)
    ra.set_(m, N=wl.N0)
    ra.time_step_(m, 0.01)
    assert np.all(np.isfinite(m.action))
    m2 = device_model(wl, np.float32)
    m2.set_doppler_velocity(np.zeros((16, 16, len(sg.kappa))), np.zeros((16, 16, len(sg.kappa))))
    with pytest.raises(ra.RippleError):  # prescribed by hand: no kappa derivative known
        m2.doppler_velocity_kappa_derivative()
    a = device_model(wl, np.float32)
    b = device_model(wl, np.float32)
    with pytest.raises(ra.RippleError):
        ra.Ensemble([a, b])
