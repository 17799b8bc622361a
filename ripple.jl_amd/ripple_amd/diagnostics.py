"""Spectral diagnostics computed on the device (Diagnostics/moments.jl,
bulk_statistics.jl).  Each returns host numpy arrays of shape (Nx, Ny) or a
float for the global integrals."""
import ctypes
import math

import numpy as np

from . import _lib


def _columns(model, which):
    Nx, Ny = model.size[0], model.size[1]
    out = [np.zeros(Nx * Ny) if w else None for w in which]
    ptrs = [_lib.dptr(a) if a is not None else None for a in out]
    _lib.check(model._ctx.lib.rpl_diagnostics(model._ctx.ptr, *ptrs))
    return [a.reshape((Nx, Ny), order="F") if a is not None else None for a in out]


def m0(model):
    """m0 = sum_mn N w (moments.jl:33-37)."""
    return _columns(model, (True, False, False))[0]


def root_mean_square_wavenumber(model):
    """sqrt(sum N (mu_xx + mu_yy) / sum N w), 0 where the action is zero (moments.jl:100-111)."""
    return _columns(model, (False, True, False))[1]


def mean_direction(model):
    """atan(sum N mu_y, sum N mu_x) (moments.jl:117-123)."""
    return _columns(model, (False, False, True))[2]


def significant_wave_height(model):
    """4 sqrt(m0) (bulk_statistics.jl:9)."""
    return 4 * np.sqrt(m0(model))


def _totals(model):
    a, e = ctypes.c_double(), ctypes.c_double()
    _lib.check(model._ctx.lib.rpl_conservation(model._ctx.ptr, ctypes.byref(a), ctypes.byref(e)))
    return a.value, e.value


def total_action(model):
    """sum_ij dx dy sum_mn N w, compensated FP64 on the device (bulk_statistics.jl:87-91)."""
    return _totals(model)[0]


def total_deep_water_energy(model, gravity=9.81):
    """sum_ij dx dy sum_mn N sqrt(g) * (radial power 1/2 measure) (bulk_statistics.jl:93-99).

    The device measure carries the model's gravity; the measure is linear in
    sqrt(g), so another `gravity` rescales the compensated total."""
    e = _totals(model)[1]
    return e if float(gravity) == model.gravity else e * math.sqrt(float(gravity) / model.gravity)
