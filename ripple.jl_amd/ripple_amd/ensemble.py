"""Batched ensembles (no counterpart in the reference, which runs one model
per process): several SpectralWaveModels of different sizes and spectral
grids advanced together, one persistent-kernel launch per RK3 stage over a
work list of every member's tiles, optionally replayed from a hipGraph, with
the launch shape picked by the library's autotuner."""
import ctypes

from . import _lib


class Ensemble:
    def __init__(self, models):
        models = list(models)
        lib = _lib.load()
        arr = (ctypes.c_void_p * len(models))(*[m._ctx.ptr.value for m in models])
        p = ctypes.c_void_p()
        _lib.check(lib.rpl_ensemble_create(ctypes.byref(p), arr, len(models)))
        self._lib, self.ptr, self.models = lib, p, models

    def close(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            self._lib.rpl_ensemble_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, dt, nsteps=1):
        """time_step! of every member, nsteps times."""
        _lib.check(self._lib.rpl_ensemble_step(self.ptr, float(dt), int(nsteps)))
        t, it = ctypes.c_double(), ctypes.c_int64()
        for m in self.models:
            _lib.check(self._lib.rpl_get_clock(m._ctx.ptr, ctypes.byref(t), ctypes.byref(it)))
            m.clock.time, m.clock.iteration = t.value, it.value
        return self

    def synchronize(self):
        _lib.check(self._lib.rpl_ensemble_synchronize(self.ptr))

    def set_launch(self, tile_rows=0, blocks=0):
        _lib.check(self._lib.rpl_ensemble_set_launch(self.ptr, int(tile_rows), int(blocks)))

    def set_split(self, min_tiles=0):
        """Members with at least `min_tiles` tiles of 64 x 4 cells x one kappa
        bin run their own stage kernel; 0 the default, negative none."""
        _lib.check(self._lib.rpl_ensemble_set_split(self.ptr, int(min_tiles)))

    def split(self):
        t, n = ctypes.c_int64(), ctypes.c_int32()
        _lib.check(self._lib.rpl_ensemble_get_split(self.ptr, ctypes.byref(t), ctypes.byref(n)))
        return {"min_tiles": t.value, "own_members": n.value}

    def launch(self):
        ty, b, t = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        _lib.check(self._lib.rpl_ensemble_get_launch(self.ptr, ctypes.byref(ty), ctypes.byref(b), ctypes.byref(t)))
        return {"tile_rows": ty.value, "blocks": b.value, "tiles": t.value, **self.split()}

    def autotune(self, dt, reps=3):
        ty, b, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        _lib.check(self._lib.rpl_ensemble_autotune(self.ptr, float(dt), int(reps), ctypes.byref(ty),
                                                   ctypes.byref(b), ctypes.byref(ms)))
        return {"tile_rows": ty.value, "blocks": b.value, "stage_ms": ms.value, **self.split()}

    def use_graph(self, enable=True):
        _lib.check(self._lib.rpl_ensemble_use_graph(self.ptr, 1 if enable else 0))
