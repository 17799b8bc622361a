"""ctypes binding of libripple_mi355x.so (the C ABI in include/ripple_mi355x.h).

The library is the only compute path: if it is missing, or no HIP device is
visible when a context is created, the calls fail loudly.  There is no CPU
fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libripple_mi355x.so")

RPL_OK = 0
RPL_F32, RPL_F64 = 0, 1
RPL_RK3, RPL_LSRK3, RPL_FORWARD_EULER = 0, 1, 2
RPL_TRANSPORT_NONE, RPL_TRANSPORT_INTRINSIC, RPL_TRANSPORT_REFRACTION = 0, 1, 2
RPL_DEPTH_INFINITE, RPL_DEPTH_CONSTANT = 0, 1
RPL_KERNEL_AUTO, RPL_KERNEL_REFERENCE, RPL_KERNEL_FAST = 0, 1, 2

ERRORS = {
    -1: "RPL_ERR_ARG",
    -2: "RPL_ERR_STATE",
    -3: "RPL_ERR_HIP",
    -4: "RPL_ERR_UNSUPPORTED",
    -5: "RPL_ERR_NOMEM",
}


class RippleError(RuntimeError):
    """A negative status returned by the C ABI."""

    def __init__(self, code, message):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code


class RippleArgumentError(RippleError, ValueError):
    """RPL_ERR_ARG: the reference raises ArgumentError in the same places."""


class rpl_config(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("nkappa", ctypes.c_int32), ("nphi", ctypes.c_int32),
        ("dtype", ctypes.c_int32), ("timestepper", ctypes.c_int32),
        ("transport", ctypes.c_int32), ("kernel", ctypes.c_int32),
    ]


class rpl_sources(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("wind_input", "drag", "wind_field", "stress_cap", "dissipation",
                                               "rotation")] + \
               [(n, ctypes.c_double) for n in (
                   "wind_speed", "wind_direction",
                   "beta_max", "z_alpha", "p_in", "alpha0", "rho_air", "rho_water", "von_karman", "gravity",
                   "cap_epsilon",
                   "vmax", "rmax", "radius", "shape", "inflow_angle", "background_x", "background_y",
                   "track_t0", "track_t1", "track_x0", "track_y0", "track_x1", "track_y1",
                   "cds", "delta1", "delta2", "p_mean",
                   "ls_cds", "ls_br", "ls_p", "ls_sigma_power",
                   "wc_rate", "wc_threshold", "wc_saturation_power", "wc_wavenumber_power",
                   "wc_reference_wavenumber")] + \
               [("nonlinear", ctypes.c_int32), ("dia_coefficient", ctypes.c_double),
                ("dia_frequency_offset", ctypes.c_double), ("dia_gravity", ctypes.c_double)]


class rpl_surrogate(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("floor", "shift", "scale", "wind_scale", "depth_feature", "rate_unit",
                                               "wind_lo", "wind_hi", "action_hi")] + \
               [("audit_stride", ctypes.c_int32), ("audit_bound", ctypes.c_float), ("terms", ctypes.c_int32),
                ("refresh", ctypes.c_int32)]


_P = ctypes.c_void_p
_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.c_int32
_I64P = ctypes.POINTER(ctypes.c_int64)
_FP = ctypes.POINTER(ctypes.c_float)

# name -> (argtypes, restype); every entry point declared in include/ripple_mi355x.h.
SIGNATURES = {
    "rpl_create": ([ctypes.POINTER(_P), ctypes.POINTER(rpl_config)], ctypes.c_int),
    "rpl_destroy": ([_P], ctypes.c_int),
    "rpl_set_geometry": ([_P, _D, _D, _DP, _DP, _DP, _DP, _I32, _D, _D], ctypes.c_int),
    "rpl_set_depth_field": ([_P, _DP], ctypes.c_int),
    "rpl_set_current": ([_P, _I32, _DP, _DP, _DP, _D], ctypes.c_int),
    "rpl_set_doppler": ([_P, _DP, _DP], ctypes.c_int),
    "rpl_get_doppler": ([_P, _DP, _DP], ctypes.c_int),
    "rpl_set_pseudomomentum_coupling": ([_P, _I32, _DP, _D], ctypes.c_int),
    "rpl_update_coupling": ([_P], ctypes.c_int),
    "rpl_set_propagation_smoothing": ([_P, _I32, _D, _D], ctypes.c_int),
    "rpl_set_topology": ([_P, _I32, _I32], ctypes.c_int),
    "rpl_get_doppler_kappa_derivative": ([_P, _DP, _DP], ctypes.c_int),
    "rpl_pseudomomentum": ([_P, _I32, _DP, _D, _I32, _DP, _DP], ctypes.c_int),
    "rpl_pseudomomentum_tendency": ([_P, _P, _D, _I32, _DP, _D, _D, _DP, _DP], ctypes.c_int),
    "rpl_set_action": ([_P, _P], ctypes.c_int),
    "rpl_get_action": ([_P, _P], ctypes.c_int),
    "rpl_tendency": ([_P, _P], ctypes.c_int),
    "rpl_time_step": ([_P, _D, _I32], ctypes.c_int),
    "rpl_synchronize": ([_P], ctypes.c_int),
    "rpl_get_clock": ([_P, _DP, _I64P], ctypes.c_int),
    "rpl_set_clock": ([_P, _D, ctypes.c_int64], ctypes.c_int),
    "rpl_set_nonfinite_check": ([_P, _I32], ctypes.c_int),
    "rpl_get_nonfinite": ([_P, _I64P, _I64P], ctypes.c_int),
    "rpl_diagnostics": ([_P, _DP, _DP, _DP], ctypes.c_int),
    "rpl_conservation": ([_P, _DP, _DP], ctypes.c_int),
    "rpl_weighted_total_noalloc": ([_P, _I32, _I32, _I32, _I32, _I32, _P, _D, _P, _P], ctypes.c_int),
    "rpl_set_sources": ([_P, ctypes.POINTER(rpl_sources), _DP, _DP], ctypes.c_int),
    "rpl_get_wind": ([_P, _DP, _DP, _DP], ctypes.c_int),
    "rpl_set_storm_track": ([_P, _I32, _DP, _DP, _DP], ctypes.c_int),
    "rpl_set_quadruplets": ([_P, _I32, ctypes.POINTER(_I32), _DP, _DP, _D, _D], ctypes.c_int),
    "rpl_set_quadruplet_rate_field": ([_P, _DP], ctypes.c_int),
    "rpl_enable_kernel_timing": ([_P, _I32], ctypes.c_int),
    "rpl_get_kernel_timing": ([_P, _DP, _I64P], ctypes.c_int),
    "rpl_get_surrogate_timing": ([_P, _DP, _I64P], ctypes.c_int),
    "rpl_get_coupling_timing": ([_P, _DP, _I64P], ctypes.c_int),
    "rpl_get_smoothing_timing": ([_P, _DP, _I64P], ctypes.c_int),
    "rpl_get_nonlinear_timing": ([_P, _DP, _I64P], ctypes.c_int),
    "rpl_get_source_timing": ([_P, _DP, _I64P], ctypes.c_int),
    "rpl_stream": ([_P], ctypes.c_void_p),
    "rpl_ensemble_create": ([ctypes.POINTER(_P), ctypes.POINTER(_P), _I32], ctypes.c_int),
    "rpl_ensemble_destroy": ([_P], ctypes.c_int),
    "rpl_ensemble_step": ([_P, _D, _I32], ctypes.c_int),
    "rpl_ensemble_synchronize": ([_P], ctypes.c_int),
    "rpl_ensemble_set_launch": ([_P, _I32, _I32], ctypes.c_int),
    "rpl_ensemble_get_launch": ([_P, ctypes.POINTER(_I32), ctypes.POINTER(_I32), _I64P], ctypes.c_int),
    "rpl_ensemble_autotune": ([_P, _D, _I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32), _DP], ctypes.c_int),
    "rpl_ensemble_set_split": ([_P, ctypes.c_int64], ctypes.c_int),
    "rpl_ensemble_get_split": ([_P, _I64P, ctypes.POINTER(_I32)], ctypes.c_int),
    "rpl_ensemble_use_graph": ([_P, _I32], ctypes.c_int),
    "rpl_mlp_create": ([ctypes.POINTER(_P), _I32, _I32, _I32, _I32, _I32, ctypes.c_uint64], ctypes.c_int),
    "rpl_mlp_destroy": ([_P], ctypes.c_int),
    "rpl_mlp_dims": ([_P, ctypes.POINTER(_I32), ctypes.POINTER(_I32), _I64P], ctypes.c_int),
    "rpl_mlp_set_params": ([_P, _FP], ctypes.c_int),
    "rpl_mlp_get_params": ([_P, _FP], ctypes.c_int),
    "rpl_mlp_forward": ([_P, _I32, _FP, _FP], ctypes.c_int),
    "rpl_mlp_gradients": ([_P, _I32, _FP, _FP, _FP, _FP], ctypes.c_int),
    "rpl_mlp_train_step": ([_P, _I32, _FP, _FP, ctypes.c_float, _FP], ctypes.c_int),
    "rpl_mlp_train_device": ([_P, ctypes.c_int64, _P, _P, _I32, _I32, ctypes.c_float, ctypes.c_float, ctypes.c_uint64,
                              _FP], ctypes.c_int),
    "rpl_attach_surrogate": ([_P, _P, ctypes.POINTER(rpl_surrogate), _FP], ctypes.c_int),
    "rpl_surrogate_stats": ([_P, _I64P, _I64P, _I64P, _I64P, _DP], ctypes.c_int),
    "rpl_surrogate_training_pairs": ([_P, ctypes.POINTER(rpl_surrogate), _FP, _FP], ctypes.c_int),
    "rpl_last_error": ([], ctypes.c_char_p),
}

_lib = None


def load():
    """Load the shared library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python ripple.jl_amd/build.py` "
            "(the HIP library is the only compute path; there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(code):
    """Raise the matching exception for a negative status."""
    if code == RPL_OK:
        return
    msg = load().rpl_last_error().decode(errors="replace")
    if code == -1:
        raise RippleArgumentError(code, msg)
    raise RippleError(code, msg)


def dptr(a):
    """Pointer to the data of a C-contiguous float64 numpy array."""
    return a.ctypes.data_as(_DP)


def vptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def fptr(a):
    """Pointer to the data of a C-contiguous float32 numpy array (None -> NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(_FP)
