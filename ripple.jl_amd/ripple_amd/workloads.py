"""Synthetic workloads: grids, currents and initial spectra of the benchmark
configurations, built deterministically from a seed (Philox, a counter-based
generator, so every consumer draws the same numbers).

Each builder returns a Workload: the constructor arguments of
SpectralWaveModel plus the initial action array and the run length.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from .grids import RectilinearGrid, PolarWaveVectorGrid, FrequencyDirectionGrid, xnodes, ynodes, PERIODIC, BOUNDED

DEFAULT_SEED = 20261203


@dataclass
class Workload:
    name: str
    grid: RectilinearGrid
    spectral_grid: object
    velocities: object  # (u, v) arrays (Nx, Ny, Nz) or None
    N0: np.ndarray      # (Nx, Ny, Nk, Nphi), Fortran order, float64
    dt: float
    nsteps: int
    model_kwargs: dict = field(default_factory=dict)

    @property
    def cell_bins(self):
        return int(np.prod(self.N0.shape))


def narrow_band_action(grid, sg, kappa0=0.4, sigma_kappa=0.05, sigma_phi=0.30):
    """exp(-((kappa - kappa0)/sigma_kappa)^2 - sin^2(phi/2)/sigma_phi^2), uniform
    in (x, y), sampled at the cell centres from (kx, ky) as the reference's
    set! passes them (examples/vortex_refraction.jl:107-111)."""
    kx = sg.kappa[:, None] * np.cos(sg.phi)[None, :]
    ky = sg.kappa[:, None] * np.sin(sg.phi)[None, :]
    kap = np.hypot(kx, ky)
    phi = np.arctan2(ky, kx)
    col = np.exp(-((kap - kappa0) / sigma_kappa) ** 2 - np.sin(phi / 2) ** 2 / sigma_phi ** 2)
    N = np.broadcast_to(col[None, None], (grid.Nx, grid.Ny) + col.shape)
    return np.asfortranarray(N, dtype=np.float64)


def vortex_refraction(float_type=np.float64, nx=64, ny=64, nz=16, nsteps=400):
    """README.md:57-108 quick start: 64x64x16 periodic box of 80 x 80, z in
    [-1, 0]; kappa = 4 bins 0.30-0.50, phi = 16 bins from 0; barotropic
    Gaussian vortex u_theta = 0.4 (r/a) exp(0.5 - 0.5 (r/a)^2), a = 10; fused
    refraction kernel, no sources, SSP-RK3, dt = 0.05."""
    Lx = Ly = 80.0
    grid = RectilinearGrid(float_type=float_type, size=(nx, ny, nz), x=(0, Lx), y=(0, Ly), z=(-1, 0),
                           halo=(3, 3, 3), topology=(PERIODIC, PERIODIC, BOUNDED))
    sg = PolarWaveVectorGrid(float_type, kappa=np.linspace(0.30, 0.50, 4),
                             phi=np.linspace(0, 2 * math.pi, 17)[:16])
    xc, yc, a, U0 = Lx / 2, Ly / 2, Lx / 8, 0.4
    x = xnodes(grid)[:, None]
    y = ynodes(grid)[None, :]
    r = np.hypot(x - xc, y - yc)
    ut = np.where(r == 0, 0.0, U0 * (r / a) * np.exp(0.5 - 0.5 * (r / a) ** 2))
    rr = np.maximum(r, np.finfo(np.float64).eps)
    u2 = -ut * (y - yc) / rr
    v2 = ut * (x - xc) / rr
    u = np.repeat(u2[:, :, None], nz, axis=2)
    v = np.repeat(v2[:, :, None], nz, axis=2)
    return Workload("vortex_refraction_64x64x16_k4_p16", grid, sg, (u, v), narrow_band_action(grid, sg),
                    0.05, nsteps, dict(horizontal_advection=None, timestepper="RK3"))


def eddy_streamfunction(nx, ny, Lx, Ly, peak_wavelength, rms_speed, seed):
    """Random-phase Gaussian-spectrum streamfunction on the periodic box; returns
    (u, v) = (-dpsi/dy, dpsi/dx) at the cell centres scaled to the rms speed."""
    rng = np.random.Generator(np.random.Philox(seed))
    kx = 2 * math.pi * np.fft.fftfreq(nx, d=Lx / nx)
    ky = 2 * math.pi * np.fft.fftfreq(ny, d=Ly / ny)
    KX, KY = np.meshgrid(kx, ky, indexing="ij")
    K = np.hypot(KX, KY)
    k0 = 2 * math.pi / peak_wavelength
    amp = np.exp(-0.5 * ((K - k0) / (0.25 * k0)) ** 2)
    amp[0, 0] = 0.0
    phase = rng.uniform(0.0, 2 * math.pi, size=K.shape)
    spec = amp * np.exp(1j * phase)
    # Cell centres sit half a cell off the FFT nodes: shift by exp(i k dx/2).
    shift = np.exp(1j * (KX * (Lx / nx) / 2 + KY * (Ly / ny) / 2))
    u = np.real(np.fft.ifft2(-1j * KY * spec * shift))
    v = np.real(np.fft.ifft2(1j * KX * spec * shift))
    rms = math.sqrt(float(np.mean(u * u + v * v)))
    return u * (rms_speed / rms), v * (rms_speed / rms)


def sheared_eddies(float_type=np.float32, nx=512, ny=512, nz=32, nkappa=32, nphi=36, nsteps=200,
                   seed=DEFAULT_SEED):
    """512^2 x 32 periodic box of 80 x 80, depth 1 (z in [-1, 0]); kappa 32 bins
    log-spaced 0.05-2.0, phi 36 bins; baroclinic eddy current (peak wavelength
    20, rms 0.3) times exp(z/0.2) at the z centres; narrow-band spectrum at
    kappa0 = 0.4; WENO5 transport through the fused refraction kernel; 200
    steps of dt = 0.02, no sources.  Deep-water intrinsic c_g, Q depth 1."""
    Lx = Ly = 80.0
    grid = RectilinearGrid(float_type=float_type, size=(nx, ny, nz), x=(0, Lx), y=(0, Ly), z=(-1, 0),
                           halo=(3, 3, 3), topology=(PERIODIC, PERIODIC, BOUNDED))
    sg = PolarWaveVectorGrid(float_type, kappa=np.geomspace(0.05, 2.0, nkappa),
                             phi=np.linspace(0, 2 * math.pi, nphi + 1)[:nphi])
    u2, v2 = eddy_streamfunction(nx, ny, Lx, Ly, 20.0, 0.3, seed)
    zc = 0.5 * (grid.z_faces[1:] + grid.z_faces[:-1])
    decay = np.exp(zc / 0.2)
    u = np.asfortranarray(u2[:, :, None] * decay[None, None, :])
    v = np.asfortranarray(v2[:, :, None] * decay[None, None, :])
    return Workload(f"sheared_eddies_{nx}x{ny}x{nz}_k{nkappa}_p{nphi}", grid, sg, (u, v),
                    narrow_band_action(grid, sg), 0.02, nsteps, dict(timestepper="RK3"))


def jonswap_action(grid, sg, Hs, fp, gravity=9.81, gamma=3.3):
    """A JONSWAP-shaped seed with cos^2 directional spreading about phi = 0,
    expressed as action density on the spectral cells and scaled so that the
    diagnostic 4 sqrt(m0) of the action equals Hs (the reference's
    significant_wave_height, Diagnostics/bulk_statistics.jl:9)."""
    f = np.sqrt(gravity * sg.kappa) / (2 * math.pi)
    sig = np.where(f <= fp, 0.07, 0.09)
    r = np.exp(-((f - fp) ** 2) / (2 * sig ** 2 * fp ** 2))
    S = f ** -5.0 * np.exp(-1.25 * (fp / f) ** 4) * gamma ** r
    # E(f) df = E(kappa) kappa dkappa, df/dkappa = g / (8 pi^2 f) for deep water.
    Ek = S * gravity / (8 * math.pi ** 2 * f) / sg.kappa
    D = np.where(np.cos(sg.phi) > 0, np.cos(sg.phi) ** 2, 0.0)
    col = Ek[:, None] * D[None, :]
    m0 = float(np.sum(col * sg.weights))
    col *= (Hs / 4) ** 2 / m0
    N = np.broadcast_to(col[None, None], (grid.Nx, grid.Ny) + col.shape)
    return np.asfortranarray(N, dtype=np.float64)


def fetch_limited_growth(float_type=np.float32, nx=256, ny=256, nz=16, nfreq=40, nphi=36, dt=10.0,
                         hours=6.0):
    """256^2 x 16 periodic box of 100 km x 100 km, depth 50 m, no current;
    FrequencyDirectionGrid of 40 log-spaced frequencies 0.05-2 Hz x 36
    directions; uniform U10 = 10 m/s along x; PrecomputedSources of
    PressureCorrelationInput (linear bulk drag, stress cap) and
    MeanSpectrumWhitecapping; JONSWAP seed Hs = 0.05 m, fp = 0.5 Hz, cos^2
    spreading; SSP-RK3 with dt = 10 s for 6 h (2160 steps); WENO5 intrinsic
    transport at the finite-depth c_g."""
    from .sources import PrecomputedSources, PressureCorrelationInput, MeanSpectrumWhitecapping, BulkWindDrag
    L = 100e3
    grid = RectilinearGrid(float_type=float_type, size=(nx, ny, nz), x=(0, L), y=(0, L), z=(-50.0, 0.0),
                           topology=(PERIODIC, PERIODIC, BOUNDED))
    sg = FrequencyDirectionGrid(float_type, frequency=np.geomspace(0.05, 2.0, nfreq),
                                phi=np.linspace(0, 2 * math.pi, nphi + 1)[:nphi])
    sources = PrecomputedSources(wind_input=PressureCorrelationInput(drag=BulkWindDrag("linear"), wind=10.0,
                                                                     direction=0.0),
                                 dissipation=MeanSpectrumWhitecapping())
    N0 = jonswap_action(grid, sg, 0.05, 0.5)
    return Workload(f"fetch_limited_growth_{nx}x{ny}x{nz}_f{nfreq}_p{nphi}", grid, sg, None, N0, dt,
                    int(round(hours * 3600 / dt)),
                    dict(timestepper="RK3", sources=sources, depth=50.0))


def wind_sea_with_dia(float_type=np.float32, nx=256, ny=256, nz=4, nfreq=40, nphi=36, dt=60.0, nsteps=10,
                      hs=2.0, fp=0.12, seed=DEFAULT_SEED):
    """The fetch-limited box (100 km x 100 km periodic, depth 50 m) with the
    full source bundle including the nonlinear transfer:
    PrecomputedSources(PressureCorrelationInput (U10 = 15 m/s along 0.3 rad,
    linear drag, stress cap), LocalSaturationDissipation,
    DiscreteInteractionApproximation(SymmetricQuadruplet())); 40 geometric
    frequencies 0.05-0.4 Hz x 36 directions (the frequency range of the
    reference's DIA tests); a developed JONSWAP sea (Hs = 2 m, fp = 0.12 Hz,
    cos^2 spreading) with +-30 % per-cell noise; WENO5 intrinsic transport,
    SSP-RK3, dt = 60 s."""
    from .sources import (PrecomputedSources, PressureCorrelationInput, LocalSaturationDissipation, BulkWindDrag,
                          DiscreteInteractionApproximation, SymmetricQuadruplet)
    L = 100e3
    grid = RectilinearGrid(float_type=float_type, size=(nx, ny, nz), x=(0, L), y=(0, L), z=(-50.0, 0.0),
                           topology=(PERIODIC, PERIODIC, BOUNDED))
    sg = FrequencyDirectionGrid(float_type, frequency=np.geomspace(0.05, 0.4, nfreq),
                                phi=np.linspace(0, 2 * math.pi, nphi + 1)[:nphi])
    sources = PrecomputedSources(
        wind_input=PressureCorrelationInput(drag=BulkWindDrag("linear"), wind=15.0, direction=0.3),
        dissipation=LocalSaturationDissipation(),
        nonlinear=DiscreteInteractionApproximation(SymmetricQuadruplet()))
    rng = np.random.Generator(np.random.Philox(seed))
    N0 = jonswap_action(grid, sg, hs, fp)
    N0 = np.asfortranarray(N0 * rng.uniform(0.7, 1.3, size=N0.shape))
    return Workload(f"wind_sea_with_dia_{nx}x{ny}_f{nfreq}_p{nphi}", grid, sg, None, N0, dt, nsteps,
                    dict(timestepper="RK3", sources=sources, depth=50.0))


def translating_hurricane(float_type=np.float32, nx=1024, ny=1024, nz=32, nkappa=32, nphi=36, dt=60.0,
                          nsteps=60):
    """1024^2 x 32 periodic box of 800 km x 800 km, depth 200 m; polar kappa
    32 bins log-spaced 0.05-2 rad/m x 36 directions; Holland hurricane (vmax
    50 m/s, rmax 30 km, B = 1.5, radius 300 km, counter-clockwise, 20 deg
    inflow) translating along x at 5 m/s from x = 200 km on the centre line;
    PrecomputedSources of PressureCorrelationInput (linear bulk drag) and
    LocalSaturationDissipation (B_r = 1.05e-2) as in the reference's
    translating-hurricane example; dt = 60 s; WENO5 intrinsic transport."""
    from .sources import (PrecomputedSources, PressureCorrelationInput, LocalSaturationDissipation, BulkWindDrag,
                          HollandHurricaneWind, LinearStormTrack)
    L = 800e3
    grid = RectilinearGrid(float_type=float_type, size=(nx, ny, nz), x=(0, L), y=(-L / 2, L / 2),
                           z=(-200.0, 0.0), topology=(PERIODIC, PERIODIC, BOUNDED))
    sg = PolarWaveVectorGrid(float_type, kappa=np.geomspace(0.05, 2.0, nkappa),
                             phi=np.linspace(0, 2 * math.pi, nphi + 1)[:nphi])
    T = 10 * 86400.0
    track = LinearStormTrack([0.0, T], [(200e3, 0.0), (200e3 + 5.0 * T, 0.0)])
    hurricane = HollandHurricaneWind(center=track, vmax=50.0, rmax=30e3, radius=300e3, shape_parameter=1.5,
                                     inflow_angle=math.radians(20.0))
    sources = PrecomputedSources(wind_input=PressureCorrelationInput(drag=BulkWindDrag("linear"), wind=hurricane),
                                 dissipation=LocalSaturationDissipation(B_r=1.05e-2))
    N0 = jonswap_action(grid, sg, 0.5, 0.15)
    return Workload(f"translating_hurricane_{nx}x{ny}x{nz}_k{nkappa}_p{nphi}", grid, sg, None, N0, dt, nsteps,
                    dict(timestepper="RK3", sources=sources, depth=200.0))


def hurricane_training_storm(index, float_type=np.float32, nx=128, ny=128, nz=8, nkappa=32, nphi=36, dt=60.0,
                             nsteps=120, seed_base=1000):
    """A randomised relative of translating_hurricane for surrogate training,
    drawn from seed seed_base + index: the same 800 km box, depth 200 m and
    32 x 36 polar grid at nx x ny cells; Holland storm of vmax uniform in
    25-60 m/s, rmax 20-60 km, radius 200-400 km, B 1.0-2.0, inflow 10-30 deg,
    counter-clockwise, starting anywhere in the box and translating at 0-10
    m/s in a random direction; JONSWAP seed of Hs 0.2-1.5 m and peak
    frequency 0.1-0.25 Hz (cos^2 about phi = 0); same source bundle."""
    from .sources import (PrecomputedSources, PressureCorrelationInput, LocalSaturationDissipation, BulkWindDrag,
                          HollandHurricaneWind, LinearStormTrack)
    rng = np.random.Generator(np.random.Philox(seed_base + index))
    L = 800e3
    grid = RectilinearGrid(float_type=float_type, size=(nx, ny, nz), x=(0, L), y=(-L / 2, L / 2),
                           z=(-200.0, 0.0), topology=(PERIODIC, PERIODIC, BOUNDED))
    sg = PolarWaveVectorGrid(float_type, kappa=np.geomspace(0.05, 2.0, nkappa),
                             phi=np.linspace(0, 2 * math.pi, nphi + 1)[:nphi])
    vmax = float(rng.uniform(25.0, 60.0))
    rmax = float(rng.uniform(20e3, 60e3))
    radius = float(rng.uniform(200e3, 400e3))
    shape = float(rng.uniform(1.0, 2.0))
    inflow = math.radians(float(rng.uniform(10.0, 30.0)))
    speed, heading = float(rng.uniform(0.0, 10.0)), float(rng.uniform(0.0, 2 * math.pi))
    x0, y0 = float(rng.uniform(0.0, L)), float(rng.uniform(-L / 2, L / 2))
    Hs, fp = float(rng.uniform(0.2, 1.5)), float(rng.uniform(0.1, 0.25))
    T = 10 * 86400.0
    track = LinearStormTrack([0.0, T], [(x0, y0), (x0 + speed * T * math.cos(heading),
                                                   y0 + speed * T * math.sin(heading))])
    wind = HollandHurricaneWind(center=track, vmax=vmax, rmax=rmax, radius=radius, shape_parameter=shape,
                                inflow_angle=inflow)
    sources = PrecomputedSources(wind_input=PressureCorrelationInput(drag=BulkWindDrag("linear"), wind=wind),
                                 dissipation=LocalSaturationDissipation(B_r=1.05e-2))
    name = (f"hurricane_training_storm_{index}_{nx}x{ny}_v{vmax:.0f}_r{rmax / 1e3:.0f}km_"
            f"u{speed:.1f}_Hs{Hs:.2f}")
    return Workload(name, grid, sg, None, jonswap_action(grid, sg, Hs, fp), dt, nsteps,
                    dict(timestepper="RK3", sources=sources, depth=200.0))


def uniform_wind_sea(index, float_type=np.float32, nx=64, ny=64, nz=8, nkappa=32, nphi=36, dt=60.0, nsteps=60,
                     wind_range=(5.0, 40.0), seed_base=2000):
    """Surrogate training sea under a uniform wind, drawn from seed seed_base +
    index: the hurricane's box, depth and 32 x 36 grid at nx x ny cells;
    U10 uniform in wind_range from a random direction; JONSWAP seed of Hs
    0.2-1.5 m, peak 0.1-0.25 Hz (cos^2 about phi = 0); same source bundle."""
    from .sources import PrecomputedSources, PressureCorrelationInput, LocalSaturationDissipation, BulkWindDrag
    rng = np.random.Generator(np.random.Philox(seed_base + index))
    L = 800e3
    grid = RectilinearGrid(float_type=float_type, size=(nx, ny, nz), x=(0, L), y=(-L / 2, L / 2),
                           z=(-200.0, 0.0), topology=(PERIODIC, PERIODIC, BOUNDED))
    sg = PolarWaveVectorGrid(float_type, kappa=np.geomspace(0.05, 2.0, nkappa),
                             phi=np.linspace(0, 2 * math.pi, nphi + 1)[:nphi])
    U = float(rng.uniform(*wind_range))
    direction = float(rng.uniform(0.0, 2 * math.pi))
    Hs, fp = float(rng.uniform(0.2, 1.5)), float(rng.uniform(0.1, 0.25))
    sources = PrecomputedSources(
        wind_input=PressureCorrelationInput(drag=BulkWindDrag("linear"), wind=U, direction=direction),
        dissipation=LocalSaturationDissipation(B_r=1.05e-2))
    return Workload(f"uniform_wind_sea_{index}_{nx}x{ny}_U{U:.1f}_Hs{Hs:.2f}", grid, sg, None,
                    jonswap_action(grid, sg, Hs, fp), dt, nsteps, dict(timestepper="RK3", sources=sources,
                                                                       depth=200.0))


def ensemble_member(index, float_type=np.float32, nz=16, sources=True, size_range=(64, 512),
                    kappa_range=(16, 40), phi_range=(24, 48), seed_base=0):
    """One member of the batched ensemble, drawn from seed seed_base + index:
    an n x n periodic box of 80 x 80 (n uniform in size_range), depth 1 with
    nz levels; kappa bins (count uniform in kappa_range) log-spaced 0.05-2,
    phi bins (count uniform in phi_range); its own eddy current (peak
    wavelength 20, rms speed uniform in 0.1-0.5) times exp(z/0.2); uniform
    wind of speed uniform in 5-25 m/s and random direction, driving
    PrecomputedSources(PressureCorrelationInput, LocalSaturationDissipation)
    when `sources`; the narrow-band spectrum at kappa0 = 0.4; dt = 0.02."""
    from .sources import PrecomputedSources, PressureCorrelationInput, LocalSaturationDissipation, BulkWindDrag
    rng = np.random.Generator(np.random.Philox(seed_base + index))
    n = int(rng.integers(size_range[0], size_range[1] + 1))
    nk = int(rng.integers(kappa_range[0], kappa_range[1] + 1))
    nphi = int(rng.integers(phi_range[0], phi_range[1] + 1))
    rms = float(rng.uniform(0.1, 0.5))
    wind = float(rng.uniform(5.0, 25.0))
    direction = float(rng.uniform(0.0, 2 * math.pi))
    Lx = Ly = 80.0
    grid = RectilinearGrid(float_type=float_type, size=(n, n, nz), x=(0, Lx), y=(0, Ly), z=(-1, 0),
                           halo=(3, 3, 3), topology=(PERIODIC, PERIODIC, BOUNDED))
    sg = PolarWaveVectorGrid(float_type, kappa=np.geomspace(0.05, 2.0, nk),
                             phi=np.linspace(0, 2 * math.pi, nphi + 1)[:nphi])
    u2, v2 = eddy_streamfunction(n, n, Lx, Ly, 20.0, rms, DEFAULT_SEED + index)
    zc = 0.5 * (grid.z_faces[1:] + grid.z_faces[:-1])
    decay = np.exp(zc / 0.2)
    u = np.asfortranarray(u2[:, :, None] * decay[None, None, :])
    v = np.asfortranarray(v2[:, :, None] * decay[None, None, :])
    kw = dict(timestepper="RK3")
    if sources:
        kw["sources"] = PrecomputedSources(
            wind_input=PressureCorrelationInput(drag=BulkWindDrag("linear"), wind=wind, direction=direction),
            dissipation=LocalSaturationDissipation())
    name = f"ensemble_member_{index}_{n}x{n}x{nz}_k{nk}_p{nphi}_U{wind:.1f}"
    return Workload(name, grid, sg, (u, v), narrow_band_action(grid, sg), 0.02, 200, kw)


def pseudomomentum_packet(float_type=np.float64, nx=64, ny=64, nz=16, nkappa=8, nphi=24, nsteps=20,
                          amplitude=5.0, depth=10.0, direction=math.pi / 6, dt=0.05):
    """A swell packet that carries its own current: an nx x ny periodic box of
    80 x 80 over z in [-depth, 0] with nz uniform levels (the Q grid), kappa
    nkappa bins log-spaced 0.1-1.0, phi nphi bins; the narrow-band spectrum at
    kappa0 = 0.4 (travelling toward phi = 0) times
    amplitude * (0.2 + exp(-|x - c|^2 / (2 * 10^2))) about the box centre c;
    the Doppler velocity is the packet's pseudomomentum projected through the
    Q kernel (PseudomomentumVelocities), recomputed every stage; WENO5
    refraction, SSP-RK3, time step dt (0.05 keeps the Courant number of the
    fastest group velocity, about 5 m/s at kappa = 0.1, at 0.2 for 64 cells;
    scale it with the cell size)."""
    from .model import PseudomomentumVelocities
    Lx = Ly = 80.0
    grid = RectilinearGrid(float_type=float_type, size=(nx, ny, nz), x=(0, Lx), y=(0, Ly), z=(-depth, 0),
                           halo=(3, 3, 3), topology=(PERIODIC, PERIODIC, BOUNDED))
    sg = PolarWaveVectorGrid(float_type, kappa=np.geomspace(0.1, 1.0, nkappa),
                             phi=np.linspace(0, 2 * math.pi, nphi + 1)[:nphi])
    x = xnodes(grid)[:, None]
    y = ynodes(grid)[None, :]
    env = amplitude * (0.2 + np.exp(-((x - Lx / 2) ** 2 + (y - Ly / 2) ** 2) / (2 * 10.0 ** 2)))
    col = np.exp(-((sg.kappa[:, None] - 0.4) / 0.1) ** 2 - np.sin((sg.phi[None, :] - direction) / 2) ** 2 / 0.3 ** 2)
    N = env[:, :, None, None] * col[None, None]
    return Workload(f"pseudomomentum_packet_{nx}x{ny}x{nz}_k{nkappa}_p{nphi}", grid, sg, None,
                    np.asfortranarray(N), dt, nsteps,
                    dict(timestepper="RK3", velocities=PseudomomentumVelocities()))
