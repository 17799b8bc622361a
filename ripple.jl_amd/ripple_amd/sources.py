"""Source terms and winds with the reference's constructors and defaults.

PressureCorrelationInput, BulkWindDrag, MeanSpectrumWhitecapping,
LocalSaturationDissipation and PrecomputedSources follow
Physics/WindInput/pressure_correlation.jl:24-52, Physics/shared/drag.jl:21-34,
Physics/Dissipation/mean_spectrum.jl:20-35, local_saturation.jl:18-28 and
Physics/Packages/precomputed_sources.jl:23-35; HollandHurricaneWind and
LinearStormTrack follow Forcing/winds.jl:4-41, 89-118.  Everything is
evaluated on the device (rpl_set_sources); these classes only carry
parameters.
"""
import ctypes
import math

import numpy as np

from . import _lib
from .grids import xnodes, ynodes


rpl_sources = _lib.rpl_sources


class BulkWindDrag:
    """`BulkWindDrag(:linear)` (Wu 1982) or `BulkWindDrag(:capped)` (Hwang 2001)."""

    def __init__(self, form="linear"):
        if form not in ("linear", "capped"):
            raise ValueError(f"unsupported drag form {form}")
        self.form = form

    def friction_velocity(self, U10):
        if self.form == "linear":
            cd = (0.8 + 0.065 * U10) * 1e-3
        else:
            cd = max((8.058 + 0.967 * U10 - 0.016 * U10 ** 2) * 1e-4, 4e-4)
        return U10 * math.sqrt(max(cd, 0.0))


class LinearStormTrack:
    """`LinearStormTrack(times, centers)` (Forcing/winds.jl:4-41): a
    piecewise-linear storm centre through the points, clamped to the ends."""

    def __init__(self, times, centers):
        t = [float(x) for x in times]
        c = [tuple(float(v) for v in p) for p in centers]
        if len(t) != len(c):
            raise ValueError("storm-track times and centers must have the same length")
        if len(t) == 0:
            raise ValueError("storm-track times and centers must be non-empty")
        if any(not t[k] > t[k - 1] for k in range(1, len(t))):
            raise ValueError("storm-track times must be strictly increasing")
        if any(len(p) != 2 for p in c):
            raise ValueError("storm-track centers must be two-component coordinates")
        self.times, self.centers = t, c

    def __call__(self, time):
        """The centre at `time` (winds.jl:28-41)."""
        t, c = self.times, self.centers
        if time <= t[0]:
            return c[0]
        if time >= t[-1]:
            return c[-1]
        upper = next(k for k in range(len(t)) if t[k] >= time)
        lower = upper - 1
        w = (time - t[lower]) / (t[upper] - t[lower])
        return ((1 - w) * c[lower][0] + w * c[upper][0], (1 - w) * c[lower][1] + w * c[upper][1])


class HollandHurricaneWind:
    """`HollandHurricaneWind(; center, vmax, rmax, radius, shape_parameter,
    inflow_angle, background, rotation)` with center a (x, y) tuple or a
    LinearStormTrack."""

    def __init__(self, center=(0.0, 0.0), vmax=1.0, rmax=1.0, radius=4.0, shape_parameter=1.5, inflow_angle=0.0,
                 background=(0.0, 0.0), rotation="counterclockwise"):
        if not vmax >= 0:
            raise ValueError("Holland hurricane vmax must be nonnegative")
        if not rmax > 0:
            raise ValueError("Holland hurricane rmax must be positive")
        if not radius > rmax:
            raise ValueError("Holland hurricane radius must exceed rmax")
        if not shape_parameter > 0:
            raise ValueError("Holland hurricane shape_parameter must be positive")
        if isinstance(center, LinearStormTrack):
            self.track = center
        else:
            self.track = LinearStormTrack([0.0, 1.0], [center, center])
        self.vmax, self.rmax, self.radius = float(vmax), float(rmax), float(radius)
        self.shape, self.inflow_angle = float(shape_parameter), float(inflow_angle)
        self.background = (float(background[0]), float(background[1]))
        self.rotation = 1 if rotation == "counterclockwise" else -1


class PressureCorrelationInput:
    def __init__(self, *, drag, wind, direction=0.0, beta_max=1.2, z_alpha=0.011, p_in=2.0, alpha0=0.0095,
                 rho_air=1.225, rho_water=1025.0, von_karman=0.4, gravity=9.81):
        self.drag, self.wind, self.direction = drag, wind, float(direction)
        self.beta_max, self.z_alpha, self.p_in, self.alpha0 = float(beta_max), float(z_alpha), float(p_in), float(alpha0)
        self.rho_air, self.rho_water = float(rho_air), float(rho_water)
        self.von_karman, self.gravity = float(von_karman), float(gravity)


class MeanSpectrumWhitecapping:
    def __init__(self, C_ds=-2.1, delta1=0.4, delta2=0.6, p=0.5, gravity=9.81):
        self.C_ds, self.delta1, self.delta2, self.p, self.gravity = (float(v) for v in (C_ds, delta1, delta2, p,
                                                                                        gravity))


class LocalSaturationDissipation:
    def __init__(self, C_ds=-2.2e-5, B_r=9.0e-4, p=2.0, sigma_power=1.0, gravity=9.81):
        self.C_ds, self.B_r, self.p, self.sigma_power, self.gravity = (float(v) for v in (C_ds, B_r, p, sigma_power,
                                                                                           gravity))


class WhitecappingDissipation:
    """`WhitecappingDissipation(; rate, saturation_threshold, saturation_power,
    wavenumber_power, reference_wavenumber)` (Sources/Sources.jl:105-122,
    583-598): S = -rate * max(m0 / threshold - 1, 0)^saturation_power *
    <kappa^wavenumber_power>_cell / reference_wavenumber^wavenumber_power * N,
    with m0 the column's zeroth moment.  The rate is a constant here."""

    def __init__(self, rate=0.0, saturation_threshold=1.0, saturation_power=1.0, wavenumber_power=1.0,
                 reference_wavenumber=1.0):
        if callable(rate) or np.ndim(rate) != 0:
            raise NotImplementedError("WhitecappingDissipation takes a constant rate on the device")
        self.rate = float(rate)
        self.saturation_threshold = float(saturation_threshold)
        self.saturation_power = float(saturation_power)
        self.wavenumber_power = float(wavenumber_power)
        self.reference_wavenumber = float(reference_wavenumber)


class SymmetricQuadruplet:
    """`SymmetricQuadruplet(; coefficient=2.78e7, frequency_offset=0.25, gravity=9.81)`
    (Physics/NonlinearInteractions/symmetric_quadruplet.jl:30-49): WAVEWATCH
    III's single-quadruplet DIA, receivers at (1 +- lambda) sigma."""

    def __init__(self, coefficient=2.78e7, frequency_offset=0.25, gravity=9.81):
        self.coefficient, self.frequency_offset, self.gravity = (float(v) for v in (coefficient, frequency_offset,
                                                                                     gravity))
        lam = self.frequency_offset
        a, b = (1 + lam) ** 2, (1 - lam) ** 2
        self.direction_offset_plus = math.acos(min(1.0, max(-1.0, (a * a - b * b + 4) / (4 * a))))
        self.direction_offset_minus = math.acos(min(1.0, max(-1.0, (b * b - a * a + 4) / (4 * b))))


class QuadrupletTransferInteraction:
    """`QuadrupletTransferInteraction(donor1, donor2, receiver1, receiver2;
    rate)` (Sources/Sources.jl:378-402): an explicit quadruplet, each bin a
    1-based (frequency m, direction n) pair as in the reference.  The rate is
    what the reference's source_value accepts (Sources.jl:1450-1467): a
    number, an (Nx, Ny) array (its [i, j] value per column), or a function of
    (x, y, t), (x, y) or (t) evaluated at the cell centres and the model
    clock."""

    def __init__(self, donor1, donor2, receiver1, receiver2, rate=0.0):
        self.bins = tuple((int(b[0]), int(b[1])) for b in (donor1, donor2, receiver1, receiver2))
        if callable(rate):
            self.rate = rate
        elif np.ndim(rate) == 0:
            self.rate = float(rate)
        elif np.ndim(rate) == 2:
            self.rate = np.array(rate, dtype=np.float64)
        else:
            raise ValueError("a quadruplet rate is a number, an (Nx, Ny) array or a function of (x, y, t), "
                             "(x, y) or (t)")


def _arity(fun):
    """How a callable source parameter is called, in the reference's order of
    `applicable` tests (Sources.jl:1458-1466): 3 for (x, y, t), 2 for (x, y),
    1 for (t)."""
    import inspect
    try:
        sig = inspect.signature(fun)
    except (TypeError, ValueError):
        return 3
    for k in (3, 2, 1):
        try:
            sig.bind(*range(k))
            return k
        except TypeError:
            pass
    raise _lib.RippleArgumentError(-1, "callable source parameters must accept (x, y, t), (x, y), or (t)")


def source_value_field(value, x_nodes, y_nodes, t):
    """source_value (Sources/Sources.jl:1450-1467) of a parameter at every
    column: (Nx, Ny) values at the cell centres and time t."""
    nx, ny = len(x_nodes), len(y_nodes)
    if callable(value):
        k = _arity(value)
        X, Y = np.meshgrid(np.asarray(x_nodes, dtype=np.float64), np.asarray(y_nodes, dtype=np.float64),
                           indexing="ij")
        args = {3: (X, Y, t), 2: (X, Y), 1: (t,)}[k]
        try:
            v = np.broadcast_to(np.asarray(value(*args), dtype=np.float64), (nx, ny))
        except Exception:  # a function of scalars only
            if k == 1:
                v = np.full((nx, ny), float(value(t)))
            else:
                v = np.array([[float(value(x, y, t) if k == 3 else value(x, y)) for y in y_nodes]
                              for x in x_nodes])
        return np.array(v, dtype=np.float64)
    if np.ndim(value) == 0:
        return np.full((nx, ny), float(value))
    a = np.asarray(value, dtype=np.float64)
    if a.shape != (nx, ny):
        raise ValueError(f"a rate array must have the grid's (Nx, Ny) = ({nx}, {ny}) shape; got {a.shape}")
    return a


class ListedRateField:
    """Uploads the per-column rates of a list whose rates are not all
    constants (rpl_set_quadruplet_rate_field), at the model clock; a rate
    that depends on time is evaluated again before every step."""

    def __init__(self, model, interactions, x_nodes, y_nodes):
        self.model, self.interactions = model, interactions
        self.x_nodes, self.y_nodes = x_nodes, y_nodes
        self.time_dependent = any(callable(q.rate) and _arity(q.rate) != 2 for q in interactions)

    def field(self, t):
        """(Nx Ny, count) rates, column i + Nx j fastest."""
        cols = [source_value_field(q.rate, self.x_nodes, self.y_nodes, t).ravel(order="F")
                for q in self.interactions]
        return np.ascontiguousarray(np.stack(cols, axis=1).ravel(order="F"))

    def upload(self, t):
        f = self.field(t)
        _lib.check(self.model._ctx.lib.rpl_set_quadruplet_rate_field(self.model._ctx.ptr, _lib.dptr(f)))


class ListedQuadruplets:
    """Sources.jl:413-417: explicit interactions with a transfer power and a
    resonance tolerance."""

    def __init__(self, interactions=(), power=1.0, resonance_tolerance=1e-12):
        self.interactions = tuple(interactions)
        self.power, self.resonance_tolerance = float(power), float(resonance_tolerance)


class DiscreteInteractionApproximation:
    """`DiscreteInteractionApproximation(method)` or
    `DiscreteInteractionApproximation(interactions; power, resonance_tolerance)`
    (Sources/Sources.jl:419-428): the nonlinear four-wave transfer with a
    SymmetricQuadruplet method or, as the reference's default, a list of
    QuadrupletTransferInteractions."""

    def __init__(self, method=(), power=1.0, resonance_tolerance=1e-12):
        if isinstance(method, (SymmetricQuadruplet, ListedQuadruplets)):
            self.method = method
        else:
            self.method = ListedQuadruplets(method, power, resonance_tolerance)


class PrecomputedSources:
    """`PrecomputedSources(; wind_input, dissipation, nonlinear)`: the bundle
    whose wind input carries the wave-supported-stress cap."""

    def __init__(self, wind_input=None, dissipation=None, nonlinear=None):
        self.wind_input, self.dissipation, self.nonlinear = wind_input, dissipation, nonlinear


def to_struct(sources):
    """rpl_sources for a bundle, or for a single term (no stress cap)."""
    s = rpl_sources()
    s.rotation = 1
    s.cap_epsilon = 0.4
    s.gravity = 9.81
    s.rho_water = 1025.0
    s.von_karman = 0.4
    s.p_mean = 0.5
    nl = None
    if isinstance(sources, PrecomputedSources):
        wi, diss, cap, nl = sources.wind_input, sources.dissipation, 1, sources.nonlinear
    elif isinstance(sources, PressureCorrelationInput):
        wi, diss, cap = sources, None, 0
    elif isinstance(sources, (MeanSpectrumWhitecapping, LocalSaturationDissipation, WhitecappingDissipation)):
        wi, diss, cap = None, sources, 0
    elif isinstance(sources, DiscreteInteractionApproximation):
        wi, diss, cap, nl = None, None, 0, sources
    else:
        raise TypeError(f"unsupported sources {type(sources).__name__}")
    if nl is not None:
        if not isinstance(nl, DiscreteInteractionApproximation):
            raise TypeError("the nonlinear member must be a DiscreteInteractionApproximation")
        q = nl.method
        if isinstance(q, ListedQuadruplets):
            # installed by attach through rpl_set_quadruplets; as a bundle
            # member its positive - damping N joins the bundle's sum through
            # the stage's additive source plane
            if not isinstance(sources, PrecomputedSources):
                return s
        else:
            s.nonlinear = 1
            s.dia_coefficient, s.dia_frequency_offset, s.dia_gravity = q.coefficient, q.frequency_offset, q.gravity
    if isinstance(diss, MeanSpectrumWhitecapping) and not isinstance(sources, PrecomputedSources):
        raise NotImplementedError("MeanSpectrumWhitecapping runs through the PrecomputedSources bundle state")
    if wi is not None:
        s.wind_input = 1
        s.drag = 0 if wi.drag.form == "linear" else 1
        s.stress_cap = cap
        for n in ("beta_max", "z_alpha", "p_in", "alpha0", "rho_air", "rho_water", "von_karman", "gravity"):
            setattr(s, n, getattr(wi, n))
        if isinstance(wi.wind, HollandHurricaneWind):
            w = wi.wind
            s.wind_field = 1
            s.vmax, s.rmax, s.radius, s.shape, s.inflow_angle = w.vmax, w.rmax, w.radius, w.shape, w.inflow_angle
            s.background_x, s.background_y = w.background
            s.rotation = w.rotation
            s.track_t0, s.track_t1 = w.track.times[0], w.track.times[-1]
            if len(w.track.times) == 1:
                s.track_t1 = s.track_t0 + 1.0  # a fixed centre
            (s.track_x0, s.track_y0), (s.track_x1, s.track_y1) = w.track.centers[0], w.track.centers[-1]
        else:
            s.wind_field = 0
            s.wind_speed = float(wi.wind)
            s.wind_direction = wi.direction
    if isinstance(diss, MeanSpectrumWhitecapping):
        s.dissipation = 1
        s.cds, s.delta1, s.delta2, s.p_mean = diss.C_ds, diss.delta1, diss.delta2, diss.p
    elif isinstance(diss, LocalSaturationDissipation):
        s.dissipation = 2
        s.ls_cds, s.ls_br, s.ls_p, s.ls_sigma_power = diss.C_ds, diss.B_r, diss.p, diss.sigma_power
    elif isinstance(diss, WhitecappingDissipation):
        s.dissipation = 3
        s.wc_rate, s.wc_threshold = diss.rate, diss.saturation_threshold
        s.wc_saturation_power, s.wc_wavenumber_power = diss.saturation_power, diss.wavenumber_power
        s.wc_reference_wavenumber = diss.reference_wavenumber
    return s


def attach(model, sources):
    """Install `sources` on the model's device context."""
    from .grids import FrequencyDirectionGrid
    lib = _lib.load()
    s = to_struct(sources)
    nl = sources.nonlinear if isinstance(sources, PrecomputedSources) else sources
    listed = nl.method if isinstance(nl, DiscreteInteractionApproximation) and \
        isinstance(nl.method, ListedQuadruplets) else None
    if (s.nonlinear or listed) and not isinstance(model.spectral_grid, FrequencyDirectionGrid):
        # symmetric_quadruplet.jl:177, Sources.jl:1275: the transfer is defined on frequency-direction grids.
        raise ValueError("DiscreteInteractionApproximation requires a FrequencyDirectionGrid")
    xn = np.ascontiguousarray(xnodes(model.grid), dtype=np.float64)
    yn = np.ascontiguousarray(ynodes(model.grid), dtype=np.float64)
    _lib.check(lib.rpl_set_quadruplets(model._ctx.ptr, 0, None, None, None, 1.0, 0.0))  # one nonlinear term at a time
    _lib.check(lib.rpl_set_sources(model._ctx.ptr, ctypes.byref(s), _lib.dptr(xn), _lib.dptr(yn)))
    wi = sources.wind_input if isinstance(sources, PrecomputedSources) else sources
    track = getattr(getattr(wi, "wind", None), "track", None)
    if track is not None and len(track.times) != 2:
        t = np.ascontiguousarray(track.times, dtype=np.float64)
        cx = np.ascontiguousarray([p[0] for p in track.centers], dtype=np.float64)
        cy = np.ascontiguousarray([p[1] for p in track.centers], dtype=np.float64)
        _lib.check(lib.rpl_set_storm_track(model._ctx.ptr, len(t), _lib.dptr(t), _lib.dptr(cx), _lib.dptr(cy)))
    else:
        _lib.check(lib.rpl_set_storm_track(model._ctx.ptr, 0, None, None, None))
    model._listed_rates = None
    if listed is not None and listed.interactions:
        bins = np.ascontiguousarray([[v - 1 for b in q.bins for v in b] for q in listed.interactions], dtype=np.int32)
        # rates given per column or as functions go in a field after the list
        rates = np.ascontiguousarray([q.rate if isinstance(q.rate, float) else 0.0 for q in listed.interactions],
                                     dtype=np.float64)
        freq = np.ascontiguousarray(model.spectral_grid.frequency, dtype=np.float64)
        _lib.check(lib.rpl_set_quadruplets(model._ctx.ptr, len(rates),
                                           bins.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _lib.dptr(rates),
                                           _lib.dptr(freq), listed.power, listed.resonance_tolerance))
        if not all(isinstance(q.rate, float) for q in listed.interactions):
            field = ListedRateField(model, listed.interactions, xn, yn)
            field.upload(model.clock.time)
            model._listed_rates = field
    return s


def wind_state(model):
    """(U10, theta_u, u_star) per column at the model clock, each (Nx, Ny)."""
    Nx, Ny = model.size[0], model.size[1]
    out = [np.zeros(Nx * Ny) for _ in range(3)]
    lib = _lib.load()
    _lib.check(lib.rpl_get_wind(model._ctx.ptr, *[_lib.dptr(a) for a in out]))
    return [a.reshape((Nx, Ny), order="F") for a in out]
