"""Physical and spectral grids, mirroring the reference constructors.

RectilinearGrid restates the parts of Oceananigans' grid the wave step reads
(uniform or given faces, centres at the face midpoints, halo, topology).
PolarWaveVectorGrid and FrequencyDirectionGrid follow
CoordinateGrids/polar_wavevector_grid.jl:18-56 and
frequency_direction_grid.jl:20-69, including their validation errors.
"""
import math

import numpy as np

PERIODIC, BOUNDED, FLAT = "Periodic", "Bounded", "Flat"


def _axis_faces(n, spec, name):
    if isinstance(spec, (tuple, list)) and len(spec) == 2 and n != 1:
        lo, hi = float(spec[0]), float(spec[1])
        return lo + (hi - lo) * np.arange(n + 1, dtype=np.float64) / n
    if isinstance(spec, (tuple, list)) and len(spec) == 2 and n == 1:
        return np.array([float(spec[0]), float(spec[1])])
    faces = np.asarray(spec, dtype=np.float64)
    if faces.shape != (n + 1,):
        raise ValueError(f"{name} faces must have length {n + 1}")
    if np.any(np.diff(faces) <= 0):
        raise ValueError(f"{name} faces must be strictly increasing")
    return faces


class RectilinearGrid:
    """`RectilinearGrid(; size, x, y, z, halo, topology)` (Oceananigans).

    x, y, z are (lo, hi) tuples for uniform spacing or arrays of faces.
    """

    def __init__(self, arch=None, float_type=np.float64, *, size, x, y, z=None, halo=(3, 3, 3),
                 topology=(PERIODIC, PERIODIC, BOUNDED)):
        if len(size) == 2:
            size = (size[0], size[1], 0)
            topology = (topology[0], topology[1], FLAT)
        self.Nx, self.Ny, self.Nz = (int(s) for s in size)
        self.float_type = np.dtype(float_type)
        self.halo = tuple(int(h) for h in halo)
        self.topology = tuple(topology)
        self.x_faces = _axis_faces(self.Nx, x, "x")
        self.y_faces = _axis_faces(self.Ny, y, "y")
        self.z_faces = _axis_faces(self.Nz, z, "z") if self.Nz > 0 else np.zeros(0)

    @property
    def Lz(self):
        return float(self.z_faces[-1] - self.z_faces[0]) if self.Nz > 0 else 0.0


def xnodes(g):
    return 0.5 * (g.x_faces[1:] + g.x_faces[:-1])


def ynodes(g):
    return 0.5 * (g.y_faces[1:] + g.y_faces[:-1])


def znodes(g):
    return 0.5 * (g.z_faces[1:] + g.z_faces[:-1])


def xfaces(g):
    return g.x_faces.copy()


def yfaces(g):
    return g.y_faces.copy()


def zfaces(g):
    return g.z_faces.copy()


def xspacings(g):
    return np.diff(g.x_faces)


def yspacings(g):
    return np.diff(g.y_faces)


def zspacings(g):
    return np.diff(g.z_faces)


def centers_to_faces(centers):
    """CoordinateGrids/spectral_grids.jl:4-21."""
    c = np.asarray(centers, dtype=np.float64)
    n = len(c)
    if n == 0:
        raise ValueError("coordinate vectors must be non-empty")
    f = np.empty(n + 1)
    if n == 1:
        f[0], f[1] = c[0] - 0.5, c[0] + 0.5
    else:
        f[1:n] = (c[:-1] + c[1:]) / 2
        f[0] = c[0] - (f[1] - c[0])
        f[n] = c[-1] + (c[-1] - f[n - 1])
    return f


def _validate_faces(name, centers, faces):
    if len(faces) != len(centers) + 1:
        raise ValueError(f"{name} faces must have length length(centers) + 1")
    if np.any(np.diff(faces) <= 0):
        raise ValueError(f"{name} faces must be strictly increasing")
    if np.any(faces[:-1] > centers) or np.any(centers > faces[1:]):
        raise ValueError(f"{name} center must lie inside its cell faces")


def _validate_centers(name, centers, positive=False):
    if len(centers) == 0:
        raise ValueError(f"{name} centers must be non-empty")
    if np.any(np.diff(centers) <= 0):
        raise ValueError(f"{name} centers must be strictly increasing")
    if positive and np.any(centers <= 0):
        raise ValueError(f"{name} centers must be positive")


def _phi_faces(phi):
    n = len(phi)
    dphi = 2 * math.pi if n == 1 else 2 * math.pi / n
    return np.array([phi[0] - dphi / 2 + i * dphi for i in range(n + 1)])


class _SpectralGrid:
    def _finish(self):
        kf, pf = self.kappa_faces, self.phi_faces
        radial = (kf[1:] ** 2 - kf[:-1] ** 2) / 2
        self.weights = radial[:, None] * np.diff(pf)[None, :]

    @property
    def shape(self):
        return (len(self.kappa), len(self.phi))

    def k_components(self, m, n):
        return self.kappa[m] * math.cos(self.phi[n]), self.kappa[m] * math.sin(self.phi[n])


class PolarWaveVectorGrid(_SpectralGrid):
    """`PolarWaveVectorGrid(; κ, φ, κ_faces, φ_faces)` (polar_wavevector_grid.jl:18-56)."""

    def __init__(self, float_type=np.float64, *, kappa=None, phi=None, kappa_faces=None, phi_faces=None):
        if kappa is None:
            raise ValueError("PolarWaveVectorGrid requires `κ`")
        if phi is None:
            raise ValueError("PolarWaveVectorGrid requires `φ`")
        self.float_type = np.dtype(float_type)
        kc = np.asarray(kappa, dtype=np.float64).ravel()
        pc = np.asarray(phi, dtype=np.float64).ravel()
        _validate_centers("φ", pc)
        kf = centers_to_faces(kc) if kappa_faces is None else np.asarray(kappa_faces, dtype=np.float64)
        if kappa_faces is None:
            kf[0] = max(kf[0], 0.0)
        pf = _phi_faces(pc) if phi_faces is None else np.asarray(phi_faces, dtype=np.float64)
        _validate_centers("κ", kc, positive=True)
        _validate_faces("κ", kc, kf)
        if np.any(kf < 0):
            raise ValueError("κ faces must be >= 0")
        _validate_faces("φ", pc, pf)
        self.kappa, self.phi, self.kappa_faces, self.phi_faces = kc, pc, kf, pf
        self.coordinate_faces_1 = kf
        self._finish()


class FrequencyDirectionGrid(_SpectralGrid):
    """`FrequencyDirectionGrid(; frequency, φ, gravity)` (frequency_direction_grid.jl:20-69):
    kappa = (2 pi f)^2 / g at centres and faces."""

    def __init__(self, float_type=np.float64, *, frequency=None, phi=None, frequency_faces=None, phi_faces=None,
                 gravity=9.81):
        if frequency is None:
            raise ValueError("FrequencyDirectionGrid requires `frequency`")
        if phi is None:
            raise ValueError("FrequencyDirectionGrid requires `φ`")
        if not gravity > 0:
            raise ValueError("gravity must be positive")
        self.float_type = np.dtype(float_type)
        fc = np.asarray(frequency, dtype=np.float64).ravel()
        pc = np.asarray(phi, dtype=np.float64).ravel()
        _validate_centers("φ", pc)
        ff = centers_to_faces(fc) if frequency_faces is None else np.asarray(frequency_faces, dtype=np.float64)
        if frequency_faces is None:
            ff[0] = max(ff[0], 0.0)
        pf = _phi_faces(pc) if phi_faces is None else np.asarray(phi_faces, dtype=np.float64)
        _validate_centers("frequency", fc, positive=True)
        _validate_faces("frequency", fc, ff)
        _validate_faces("φ", pc, pf)
        self.gravity = float(gravity)
        self.frequency, self.frequency_faces = fc, ff
        self.kappa = (2 * math.pi * fc) ** 2 / gravity
        self.kappa_faces = (2 * math.pi * ff) ** 2 / gravity
        self.phi, self.phi_faces = pc, pf
        self.coordinate_faces_1 = ff
        self._finish()


def spectral_cell_measures(g):
    return g.weights.copy()
