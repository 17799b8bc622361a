"""SpectralWaveModel, set_, time_step_, Simulation and run_ on the MI355X.

Mirrors Models/spectral_wave_model.jl:155-211 (construction and validation),
Models/tendencies.jl:18-109 (which fused tendency runs), ProductFields/set.jl
(set!), Models/time_step.jl:119-170 (time_step!) and Oceananigans'
Simulation/run!.  Julia's `f!` names are spelt `f_` here.  All state lives in
device memory owned by one C-ABI context; this module only marshals.
"""
import ctypes
import math

import numpy as np

from . import _lib
from .grids import PERIODIC, BOUNDED, RectilinearGrid, PolarWaveVectorGrid, FrequencyDirectionGrid, xnodes, ynodes


class InfiniteDepth:
    """Deep-water marker (the reference's default `depth=InfiniteDepth()`)."""

    def __repr__(self):
        return "InfiniteDepth()"


class HorizontalField:
    """Values on the horizontal wave grid, (Nx, Ny): the stand-in for the
    Oceananigans Field a reference user passes as a spatially varying
    `depth` (lagrangian_velocities.jl:128-147)."""

    def __init__(self, values):
        self.data = np.array(values, dtype=np.float64)
        if self.data.ndim != 2:
            raise ValueError("a HorizontalField holds an (Nx, Ny) array")


WENO = "WENO"
_TIMESTEPPERS = {"RK3": _lib.RPL_RK3, "LSRK3": _lib.RPL_LSRK3, "LowStorageRK3": _lib.RPL_LSRK3,
                 "ForwardEuler": _lib.RPL_FORWARD_EULER}
_KERNELS = {"auto": _lib.RPL_KERNEL_AUTO, "reference": _lib.RPL_KERNEL_REFERENCE, "fast": _lib.RPL_KERNEL_FAST}


class PseudomomentumVelocities:
    """`PseudomomentumVelocities(; q_grid, vertical_grid, Nz, surface_spacing,
    vertical_stretching)` (Coupling/lagrangian_velocities.jl:65-81): the
    current is the wave field's own pseudomomentum, recomputed from the action
    before every tendency evaluation.  The Q grid is `q_grid` (or
    `vertical_grid`), else the model grid when it has vertical cells, else an
    automatic stretched grid over the finite model depth."""

    def __init__(self, *, q_grid=None, vertical_grid=None, depth=None, Nz=None, surface_spacing=None,
                 vertical_stretching=2):
        if depth is not None:
            raise ValueError("`depth` is a SpectralWaveModel kwarg; pass `SpectralWaveModel(...; depth=...)` "
                             "instead of putting `depth` in `velocities`")
        if q_grid is not None and vertical_grid is not None:
            raise ValueError("PseudomomentumVelocities cannot contain both `q_grid` and `vertical_grid`")
        self.q_grid = q_grid if q_grid is not None else vertical_grid
        self.Nz = Nz
        self.surface_spacing = surface_spacing
        self.vertical_stretching = vertical_stretching


def automatic_q_grid(grid, spectral_grid, depth, Nz=None, surface_spacing=None, vertical_stretching=2):
    """lagrangian_velocities.jl:196-225: Nz faces z_k = -H (1 - k/Nz)^stretch
    over the finite depth H, with Nz = max(16, ceil((H / dz_surface)^(1/stretch)))
    and dz_surface = 1 / (20 kappa_max) unless given."""
    if not isinstance(spectral_grid, PolarWaveVectorGrid):
        raise ValueError("automatic Q-grid selection requires a PolarWaveVectorGrid")
    if depth is None or isinstance(depth, InfiniteDepth):
        raise ValueError("automatic Q-grid selection requires finite model `depth`")
    # lagrangian_velocities.jl:199: the deepest column (max_depth)
    H = float(np.max(depth.data)) if isinstance(depth, HorizontalField) else float(depth)
    dz_surface = 1.0 / (20 * float(np.max(spectral_grid.kappa))) if surface_spacing is None else float(surface_spacing)
    if not dz_surface > 0:
        raise ValueError("automatic Q-grid `surface_spacing` must be positive")
    stretch = float(vertical_stretching)
    if not stretch > 1:
        raise ValueError("automatic Q-grid `vertical_stretching` must be greater than 1")
    Nz = max(16, math.ceil((H / dz_surface) ** (1 / stretch))) if Nz is None else int(Nz)
    if Nz <= 0:
        raise ValueError("automatic Q-grid `Nz` must be positive")
    z = np.array([-H * (1 - k / Nz) ** stretch for k in range(Nz + 1)])
    return RectilinearGrid(float_type=grid.float_type, size=(grid.Nx, grid.Ny, Nz),
                           x=np.asarray(grid.x_faces), y=np.asarray(grid.y_faces), z=z,
                           topology=(grid.topology[0], grid.topology[1], BOUNDED))


class SpatialAveraging:
    """`SpatialAveraging(; αs=0.5, αn=0.5)` (Models/propagation_smoothing.jl:22-51):
    Tolman's garden-sprinkler alleviation, applied to the action after every
    time step."""

    def __init__(self, alpha_s=0.5, alpha_n=0.5):
        self.alpha_s = float(alpha_s)
        self.alpha_n = float(alpha_n)


class Clock:
    def __init__(self, time=0.0, iteration=0):
        self.time = float(time)
        self.iteration = int(iteration)


class _Context:
    """Owns one rpl_ctx; destroyed with the model."""

    def __init__(self, cfg):
        self.lib = _lib.load()
        self.ptr = ctypes.c_void_p()
        _lib.check(self.lib.rpl_create(ctypes.byref(self.ptr), ctypes.byref(cfg)))

    def __del__(self):
        ptr = getattr(self, "ptr", None)
        if ptr is not None and ptr.value:
            self.lib.rpl_destroy(ptr)
            self.ptr = ctypes.c_void_p()


def _evaluate_on_centres(fun, grid, nz_nodes=None):
    x = xnodes(grid)[:, None, None]
    y = ynodes(grid)[None, :, None]
    z = nz_nodes[None, None, :]
    try:
        v = np.asarray(fun(x, y, z), dtype=np.float64)
        return np.broadcast_to(v, (len(x.ravel()), len(y.ravel()), len(z.ravel()))).copy()
    except Exception:  # scalar-only user function
        return np.vectorize(fun, otypes=[np.float64])(x, y, z)


class SpectralWaveModel:
    """`SpectralWaveModel(grid, spectral_grid; velocities, horizontal_advection,
    spectral_advection, sources, depth, timestepper, clock)`.

    velocities: dict/tuple with `u`, `v` as (Nx, Ny, Nz) arrays (or callables of
    (x, y, z) sampled at cell centres), the prescribed Lagrangian-mean current.
    `kernel` picks the stencil implementation ("auto", "reference", "fast").
    """

    def __init__(self, grid, spectral_grid, *, velocities=None, horizontal_advection=WENO,
                 spectral_advection=WENO, sources=None, depth=None, timestepper="ForwardEuler",
                 clock=None, kernel="auto", gravity=9.81, propagation_smoothing=None):
        if not isinstance(grid, RectilinearGrid):
            raise TypeError(f"grid must be a RectilinearGrid; got {type(grid).__name__}")
        if not isinstance(spectral_grid, (PolarWaveVectorGrid, FrequencyDirectionGrid)):
            raise TypeError("spectral_grid must be a PolarWaveVectorGrid or FrequencyDirectionGrid")
        if timestepper not in _TIMESTEPPERS:
            raise ValueError(f"unsupported timestepper {timestepper}")
        if any(t not in (PERIODIC, BOUNDED) for t in grid.topology[:2]):
            raise NotImplementedError("x and y must be Periodic or Bounded")
        self.grid, self.spectral_grid = grid, spectral_grid
        # lagrangian_velocities.jl:174-189 (validate_model_depth): InfiniteDepth,
        # a positive number, a function of (x, y) or a Field on the horizontal
        # grid; a function or Field is a depth per column
        self.depth_field = None
        if callable(depth) and not isinstance(depth, InfiniteDepth):
            X, Y = np.meshgrid(xnodes(grid), ynodes(grid), indexing="ij")
            try:
                vals = np.broadcast_to(np.asarray(depth(X, Y), dtype=np.float64), (grid.Nx, grid.Ny))
            except Exception:  # a function of scalars only
                vals = np.array([[float(depth(x, y)) for y in ynodes(grid)] for x in xnodes(grid)])
            self.depth_field = np.array(vals, dtype=np.float64)
        elif isinstance(depth, HorizontalField):
            if depth.data.shape != (grid.Nx, grid.Ny):
                raise _lib.RippleArgumentError(-1, f"model `depth` Field must have horizontal size "
                                                   f"{(grid.Nx, grid.Ny)}")
            self.depth_field = depth.data.copy()
        elif depth is not None and not isinstance(depth, InfiniteDepth) and np.ndim(depth) != 0:
            raise _lib.RippleArgumentError(-1, "model `depth` must be InfiniteDepth(), a positive number, a function, "
                                               "or a Field on the horizontal wave grid; raw arrays are not accepted")
        if self.depth_field is not None:
            if not (self.depth_field > 0).all():
                raise _lib.RippleArgumentError(-1, "model `depth` values must be positive")
            self.depth = HorizontalField(self.depth_field)
        else:
            self.depth = InfiniteDepth() if depth is None else depth
            if not isinstance(self.depth, InfiniteDepth) and not float(self.depth) > 0:
                raise ValueError("depth must be positive")
        self.horizontal_advection = horizontal_advection
        self.spectral_advection = spectral_advection
        self.sources = sources
        self.timestepper = timestepper
        self.clock = clock if clock is not None else Clock()
        self.gravity = float(gravity)
        self.dtype = np.dtype(np.float64 if grid.float_type == np.float64 or
                              spectral_grid.float_type == np.float64 else np.float32)
        Nk, Np = spectral_grid.shape
        self.size = (grid.Nx, grid.Ny, Nk, Np)

        has_coupling = velocities is not None
        self.pseudomomentum = isinstance(velocities, PseudomomentumVelocities)
        if has_coupling and not isinstance(spectral_grid, PolarWaveVectorGrid):
            raise ValueError("CWCM current coupling requires a PolarWaveVectorGrid")
        # Models/tendencies.jl:18-109.
        if has_coupling and spectral_advection is not None:
            transport = _lib.RPL_TRANSPORT_REFRACTION
        elif horizontal_advection is None:
            transport = _lib.RPL_TRANSPORT_NONE
        elif horizontal_advection == WENO and not has_coupling and Nk * Np > 1:
            # tendencies.jl:45-50: the fused kernel on Periodic x Periodic grids
            # only; the reference takes Oceananigans' per-bin advection otherwise
            if grid.topology[0] != PERIODIC or grid.topology[1] != PERIODIC:
                raise NotImplementedError("intrinsic WENO transport is built for Periodic x Periodic grids only")
            transport = _lib.RPL_TRANSPORT_INTRINSIC
        else:
            raise NotImplementedError("only the fused transport kernels are built (WENO, periodic, >1 bin)")
        self.transport = transport

        cfg = _lib.rpl_config(grid.Nx, grid.Ny, Nk, Np,
                              _lib.RPL_F64 if self.dtype == np.float64 else _lib.RPL_F32,
                              _TIMESTEPPERS[timestepper], transport, _KERNELS[kernel])
        self._ctx = _Context(cfg)
        lib, ctx = self._ctx.lib, self._ctx.ptr
        sg = spectral_grid
        depth_kind = _lib.RPL_DEPTH_INFINITE if isinstance(self.depth, InfiniteDepth) else _lib.RPL_DEPTH_CONSTANT
        # a depth field replaces the constant depth's tables (rpl_set_depth_field below)
        geometry_depth = float(self.depth_field.mean()) if self.depth_field is not None else self.depth
        arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in
                  (sg.kappa, sg.kappa_faces, sg.phi, sg.phi_faces)]
        _lib.check(lib.rpl_set_geometry(ctx, float(grid.x_faces[1] - grid.x_faces[0]),
                                        float(grid.y_faces[1] - grid.y_faces[0]),
                                        *[_lib.dptr(a) for a in arrays], depth_kind,
                                        0.0 if depth_kind == _lib.RPL_DEPTH_INFINITE else float(geometry_depth),
                                        float(gravity)))
        if self.depth_field is not None:
            # transport.jl:102-139: the group velocity of every column at its depth
            df = np.ascontiguousarray(self.depth_field.ravel(order="F"))
            _lib.check(lib.rpl_set_depth_field(ctx, _lib.dptr(df)))
        self.velocities = None
        self.q_grid = None
        if self.pseudomomentum:
            self._set_pseudomomentum(velocities)
        elif has_coupling:
            self._set_velocities(velocities)
        _lib.check(lib.rpl_set_clock(ctx, self.clock.time, self.clock.iteration))
        # the refraction kernel wraps x and y on either topology
        # (wave_current_refraction.jl:97-111); SpatialAveraging clamps on a
        # bounded axis (propagation_smoothing.jl:97-100)
        _lib.check(lib.rpl_set_topology(ctx, int(grid.topology[0] == PERIODIC), int(grid.topology[1] == PERIODIC)))
        self.propagation_smoothing = propagation_smoothing
        if propagation_smoothing is not None:
            if not isinstance(propagation_smoothing, SpatialAveraging):
                raise TypeError("propagation_smoothing must be a SpatialAveraging (or None)")
            _lib.check(lib.rpl_set_propagation_smoothing(ctx, 1, propagation_smoothing.alpha_s,
                                                         propagation_smoothing.alpha_n))
        if sources is not None:
            from .sources import attach
            attach(self, sources)

    # --- coupling -----------------------------------------------------------
    def _set_velocities(self, velocities):
        g = self.grid
        if g.Nz < 1:
            raise ValueError("QTransform requires a grid with a resolved vertical coordinate")
        u, v = (velocities["u"], velocities["v"]) if isinstance(velocities, dict) else velocities
        zc = 0.5 * (g.z_faces[1:] + g.z_faces[:-1])
        fields = []
        for f in (u, v):
            a = _evaluate_on_centres(f, g, zc) if callable(f) else np.asarray(f, dtype=np.float64)
            if a.shape != (g.Nx, g.Ny, g.Nz):
                raise ValueError(f"velocity arrays must have size {(g.Nx, g.Ny, g.Nz)}")
            fields.append(np.asfortranarray(a))
        # Coupling/lagrangian_velocities.jl:113-115: Q depth is the column depth
        # for InfiniteDepth, else the model depth.
        # (a depth field is each column's Q depth: the library takes it from
        # rpl_set_depth_field and ignores this value)
        q_depth = g.Lz if isinstance(self.depth, InfiniteDepth) else (
            float(self.depth_field.mean()) if self.depth_field is not None else float(self.depth))
        zf = np.ascontiguousarray(g.z_faces, dtype=np.float64)
        uf = np.ascontiguousarray(fields[0].ravel(order="F"))
        vf = np.ascontiguousarray(fields[1].ravel(order="F"))
        _lib.check(self._ctx.lib.rpl_set_current(self._ctx.ptr, g.Nz, _lib.dptr(zf), _lib.dptr(uf),
                                                 _lib.dptr(vf), q_depth))
        self.velocities = (fields[0], fields[1])

    def _q_depth(self, q_grid):
        # lagrangian_velocities.jl:113-115: the Q grid's column depth for
        # InfiniteDepth, else the model depth.
        if self.depth_field is not None:
            return 0.0  # the library takes each column's depth from rpl_set_depth_field
        return q_grid.Lz if isinstance(self.depth, InfiniteDepth) else float(self.depth)

    def _set_pseudomomentum(self, pv):
        """build_coupling(::PseudomomentumVelocities) (lagrangian_velocities.jl:294-311)."""
        g = self.grid
        if pv.q_grid is not None:
            q_grid = pv.q_grid
            if (q_grid.Nx, q_grid.Ny) != (g.Nx, g.Ny) or not np.array_equal(q_grid.x_faces, g.x_faces) or \
                    not np.array_equal(q_grid.y_faces, g.y_faces):
                raise ValueError("the Q grid must share the model grid's horizontal cells")
        elif g.Nz < 1:
            if isinstance(self.depth, InfiniteDepth):
                raise ValueError("Flat wave grids require finite model `depth` or `q_grid` for "
                                 "PseudomomentumVelocities")
            q_grid = automatic_q_grid(g, self.spectral_grid, self.depth, pv.Nz, pv.surface_spacing,
                                      pv.vertical_stretching)
        else:
            q_grid = g
        if q_grid.Nz < 1:
            raise ValueError("QTransform requires a grid with a resolved vertical coordinate")
        zf = np.ascontiguousarray(q_grid.z_faces, dtype=np.float64)
        _lib.check(self._ctx.lib.rpl_set_pseudomomentum_coupling(self._ctx.ptr, q_grid.Nz, _lib.dptr(zf),
                                                                 self._q_depth(q_grid)))
        self.q_grid = q_grid
        self.velocities = pv

    def doppler_velocity(self):
        """(Ux, Uy) caches, each (Nx, Ny, Nkappa)."""
        Nx, Ny, Nk, _ = self.size
        Ux = np.zeros(Nx * Ny * Nk)
        Uy = np.zeros(Nx * Ny * Nk)
        _lib.check(self._ctx.lib.rpl_get_doppler(self._ctx.ptr, _lib.dptr(Ux), _lib.dptr(Uy)))
        return Ux.reshape((Nx, Ny, Nk), order="F"), Uy.reshape((Nx, Ny, Nk), order="F")

    def doppler_velocity_kappa_derivative(self):
        """(dUx/dkappa, dUy/dkappa) caches, each (Nx, Ny, Nkappa)."""
        Nx, Ny, Nk, _ = self.size
        a, b = np.zeros(Nx * Ny * Nk), np.zeros(Nx * Ny * Nk)
        _lib.check(self._ctx.lib.rpl_get_doppler_kappa_derivative(self._ctx.ptr, _lib.dptr(a), _lib.dptr(b)))
        return a.reshape((Nx, Ny, Nk), order="F"), b.reshape((Nx, Ny, Nk), order="F")

    def set_doppler_velocity(self, Ux, Uy):
        a = [np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel(order="F")) for x in (Ux, Uy)]
        _lib.check(self._ctx.lib.rpl_set_doppler(self._ctx.ptr, _lib.dptr(a[0]), _lib.dptr(a[1])))

    # --- state ----------------------------------------------------------------
    @property
    def action(self):
        """Interior of the wave-action ProductField as an (Nx, Ny, Nk, Nphi) array."""
        return interior(self)

    def _upload(self, values):
        a = np.asarray(values, dtype=self.dtype)
        if a.shape != self.size:
            raise ValueError(f"array size {a.shape} does not match field size {self.size}")
        flat = np.ascontiguousarray(a.ravel(order="F"))
        _lib.check(self._ctx.lib.rpl_set_action(self._ctx.ptr, _lib.vptr(flat)))

    def tendency(self):
        """compute_tendencies!(model) followed by interior(model.tendencies)."""
        rates = getattr(self, "_listed_rates", None)
        if rates is not None and rates.time_dependent:
            rates.upload(self.clock.time)
        G = np.empty(int(np.prod(self.size)), dtype=self.dtype)
        _lib.check(self._ctx.lib.rpl_tendency(self._ctx.ptr, _lib.vptr(G)))
        return G.reshape(self.size, order="F")

    def synchronize(self):
        _lib.check(self._ctx.lib.rpl_synchronize(self._ctx.ptr))

    def set_nonfinite_check(self, interval):
        """Check the action for NaN and infinite values after every
        interval-th step (0: off); clears the record."""
        _lib.check(self._ctx.lib.rpl_set_nonfinite_check(self._ctx.ptr, int(interval)))

    def nonfinite(self):
        """(first checked iteration with a non-finite value or -1, their count then)."""
        it, n = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self._ctx.lib.rpl_get_nonfinite(self._ctx.ptr, ctypes.byref(it), ctypes.byref(n)))
        return it.value, n.value

    def set_clock(self, time, iteration):
        """model.clock = (time, iteration): with set_(model, N=...) the state
        a run resumes from."""
        _lib.check(self._ctx.lib.rpl_set_clock(self._ctx.ptr, float(time), int(iteration)))
        self.clock.time, self.clock.iteration = float(time), int(iteration)

    @property
    def stream(self):
        return self._ctx.lib.rpl_stream(self._ctx.ptr)

    @property
    def context(self):
        return self._ctx.ptr


def interior(model):
    """interior(model.action): host copy, (Nx, Ny, Nk, Nphi), Fortran order."""
    N = np.empty(int(np.prod(model.size)), dtype=model.dtype)
    _lib.check(model._ctx.lib.rpl_get_action(model._ctx.ptr, _lib.vptr(N)))
    return N.reshape(model.size, order="F")


def set_(model, N=None):
    """`set!(model; N)`: N is a number, an (Nx, Ny, Nk, Nphi) array, or a
    function of (x, y, kx, ky) sampled at cell centres (ProductFields/set.jl:5-48)."""
    if N is None:
        return model
    Nx, Ny, Nk, Np = model.size
    if np.isscalar(N):
        model._upload(np.full(model.size, N, dtype=model.dtype))
    elif callable(N):
        sg = model.spectral_grid
        kx = sg.kappa[:, None] * np.cos(sg.phi)[None, :]
        ky = sg.kappa[:, None] * np.sin(sg.phi)[None, :]
        x = xnodes(model.grid)[:, None, None, None]
        y = ynodes(model.grid)[None, :, None, None]
        try:
            vals = np.asarray(N(x, y, kx[None, None], ky[None, None]), dtype=np.float64)
            vals = np.broadcast_to(vals, model.size)
        except Exception:
            vals = np.vectorize(N, otypes=[np.float64])(x, y, kx[None, None], ky[None, None])
        model._upload(vals)
    else:
        model._upload(N)
    return model


def _read_clock(model):
    t = ctypes.c_double()
    it = ctypes.c_int64()
    _lib.check(model._ctx.lib.rpl_get_clock(model._ctx.ptr, ctypes.byref(t), ctypes.byref(it)))
    model.clock.time, model.clock.iteration = t.value, it.value


def time_step_(model, dt, nsteps=1):
    """`time_step!(model, dt)`, nsteps times, on the device.  Listed
    quadruplet rates that depend on time are evaluated at the clock before
    each step (every stage of a step sees the step's start time, as
    model.clock.time in the reference)."""
    if not dt > 0:
        raise _lib.RippleArgumentError(-1, "time step must be positive")
    rates = getattr(model, "_listed_rates", None)
    if rates is not None and rates.time_dependent:
        for _ in range(int(nsteps)):
            rates.upload(model.clock.time)
            _lib.check(model._ctx.lib.rpl_time_step(model._ctx.ptr, float(dt), 1))
            _read_clock(model)
        return model
    _lib.check(model._ctx.lib.rpl_time_step(model._ctx.ptr, float(dt), int(nsteps)))
    _read_clock(model)
    return model


def update_coupling_(model):
    """`update_coupling!(model)` (Coupling/current_coupling.jl:123-148): the
    pseudomomentum coupling's U and dU/dkappa from the current action; a no-op
    for a prescribed current."""
    _lib.check(model._ctx.lib.rpl_update_coupling(model._ctx.ptr))
    return model


def _pseudomomentum(model, depth, q_grid, cell_average):
    q_grid = q_grid if q_grid is not None else (model.q_grid if model.q_grid is not None else model.grid)
    if q_grid.Nz < 1:
        raise ValueError("QTransform requires a grid with a resolved vertical coordinate")
    d = model._q_depth(q_grid) if depth is None else float(depth)
    zf = np.ascontiguousarray(q_grid.z_faces, dtype=np.float64)
    Nx, Ny = model.size[:2]
    px, py = np.zeros(Nx * Ny * q_grid.Nz), np.zeros(Nx * Ny * q_grid.Nz)
    _lib.check(model._ctx.lib.rpl_pseudomomentum(model._ctx.ptr, q_grid.Nz, _lib.dptr(zf), d, int(cell_average),
                                                 _lib.dptr(px), _lib.dptr(py)))
    shp = (Nx, Ny, q_grid.Nz)
    return px.reshape(shp, order="F"), py.reshape(shp, order="F")


def compute_pseudomomentum_cell_averages(model, depth=None, q_grid=None):
    """compute_pseudomomentum_cell_averages(N, depth, qtransform)
    (Coupling/pseudomomentum.jl:200-210) of the model's action on the Q grid
    (default: the coupling's, else the model grid) and depth (default: the
    coupling's projection depth): (px, py), each (Nx, Ny, Nz)."""
    return _pseudomomentum(model, depth, q_grid, True)


def compute_pseudomomentum_cell_integrals(model, depth=None, q_grid=None):
    """compute_pseudomomentum_cell_integrals (pseudomomentum.jl:189-199):
    the cell averages times |dz_k|."""
    return _pseudomomentum(model, depth, q_grid, False)


pseudomomentum_fields = compute_pseudomomentum_cell_averages


def compute_pseudomomentum_tendency_cell_averages(model, old_N, dt, depth=None, q_grid=None, coefficient=1.0):
    """compute_pseudomomentum_tendency_cell_averages!(ptx, pty, new_N, old_N, dt,
    depth, qtransform) (Coupling/pseudomomentum.jl:219-252) with new_N the
    model's action: the cell-averaged pseudomomentum of (new_N - old_N) / dt,
    times `coefficient`."""
    if not dt > 0:
        raise _lib.RippleArgumentError(-1, "pseudomomentum tendency timestep must be positive")
    q_grid = q_grid if q_grid is not None else (model.q_grid if model.q_grid is not None else model.grid)
    if q_grid.Nz < 1:
        raise ValueError("QTransform requires a grid with a resolved vertical coordinate")
    old = np.asarray(old_N, dtype=model.dtype)
    if old.shape != model.size:
        raise ValueError("new and old wave-action fields must have matching sizes")
    d = model._q_depth(q_grid) if depth is None else float(depth)
    zf = np.ascontiguousarray(q_grid.z_faces, dtype=np.float64)
    flat = np.ascontiguousarray(old.ravel(order="F"))
    Nx, Ny = model.size[:2]
    px, py = np.zeros(Nx * Ny * q_grid.Nz), np.zeros(Nx * Ny * q_grid.Nz)
    _lib.check(model._ctx.lib.rpl_pseudomomentum_tendency(model._ctx.ptr, _lib.vptr(flat), float(dt), q_grid.Nz,
                                                          _lib.dptr(zf), d, float(coefficient), _lib.dptr(px),
                                                          _lib.dptr(py)))
    shp = (Nx, Ny, q_grid.Nz)
    return px.reshape(shp, order="F"), py.reshape(shp, order="F")


pseudomomentum_tendency_fields = compute_pseudomomentum_tendency_cell_averages


def cwcm_momentum_tendency_fields(model, old_N, dt, depth=None, q_grid=None, coefficient=-1.0):
    """cwcm_momentum_tendency_fields! (pseudomomentum.jl:254-270): the
    pseudomomentum tendency scaled by coefficient (default -1), the momentum
    forcing a current model receives."""
    return compute_pseudomomentum_tendency_cell_averages(model, old_N, dt, depth, q_grid, coefficient)


class Simulation:
    """`Simulation(model; Δt, stop_iteration, stop_time)` (Oceananigans), with
    `output_writers`, a dict of writers (output.JLD2Writer) that `run_` calls
    on their schedules."""

    def __init__(self, model, *, dt, stop_iteration=None, stop_time=None):
        if not dt > 0:
            raise ValueError("Δt must be positive")
        self.model, self.dt = model, float(dt)
        self.stop_iteration, self.stop_time = stop_iteration, stop_time
        self.output_writers = {}


def run_(sim):
    """`run!(simulation)`: advance until stop_iteration / stop_time.  The
    output writers write at the starting iteration and after every step
    their schedule actuates on; the steps between two writes run as one
    device call."""
    m = sim.model
    if sim.stop_iteration is not None:
        n = int(sim.stop_iteration) - m.clock.iteration
    elif sim.stop_time is not None:
        n = int(math.ceil((sim.stop_time - m.clock.time) / sim.dt - 1e-9))
    else:
        raise ValueError("Simulation needs stop_iteration or stop_time")
    writers = list(sim.output_writers.values())
    try:
        for w in writers:
            if w.schedule(m):
                w.write(m)
        while n > 0:
            k = min([n] + [w.schedule.steps_to_next(m, sim.dt) for w in writers])
            time_step_(m, sim.dt, k)
            n -= k
            for w in writers:
                if w.schedule(m):
                    w.write(m)
    finally:
        for w in writers:
            w.close()
    m.synchronize()
    return sim
