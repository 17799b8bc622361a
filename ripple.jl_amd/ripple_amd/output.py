"""Output writers of a Simulation: `JLD2Writer` with `IterationInterval` and
`TimeInterval` schedules (Oceananigans' output machinery as Ripple uses it).

The reference writes its wave action through Oceananigans' JLD2Writer, with
one hook of its own: `fetch_output(field::ProductField, model) =
interior(field)` (src/OceananigansIntegration.jl:7-10), so the file holds the
interior (Nx, Ny, Nkappa, Nphi) array; test/integration/model_api.jl:49-60
runs one iteration of a Simulation with `JLD2Writer(model, (; N =
model.action); filename, schedule = IterationInterval(1))`.

Here the snapshot is downloaded from the device (rpl_get_action) when the
schedule actuates and written with the layout of Oceananigans' writer:
`timeseries/<name>/<iteration>` for each output and `timeseries/t/<iteration>`
for the clock time, plus the grid's sizes and faces under `grid/`.  The
file is a JLD2 file (jld2.JLD2File): Julia's JLD2 reads the arrays back with
their Julia dimensions.
"""
import math

import numpy as np

from .jld2 import JLD2File


class IterationInterval:
    """Actuates every `interval` iterations, iteration 0 included."""

    def __init__(self, interval):
        self.interval = int(interval)
        if self.interval < 1:
            raise ValueError("IterationInterval needs a positive interval")

    def __call__(self, model):
        return model.clock.iteration % self.interval == 0

    def steps_to_next(self, model, dt):
        r = model.clock.iteration % self.interval
        return self.interval - r


class TimeInterval:
    """Actuates when the clock reaches the next multiple of `interval` (and at
    the time the simulation starts)."""

    def __init__(self, interval):
        self.interval = float(interval)
        if not self.interval > 0:
            raise ValueError("TimeInterval needs a positive interval")
        self.next_time = None

    def __call__(self, model):
        t = model.clock.time
        if self.next_time is None:
            self.next_time = t
        if t >= self.next_time - 1e-12 * max(1.0, abs(t)):
            self.next_time = (math.floor(t / self.interval + 1e-9) + 1) * self.interval
            return True
        return False

    def steps_to_next(self, model, dt):
        if self.next_time is None:
            return 1
        return max(1, int(math.ceil((self.next_time - model.clock.time) / dt - 1e-9)))


def fetch_output(output, model):
    """The array an output stands for at the model's current state:
    "N" (or "action") is the interior wave action, the reference's
    fetch_output of a ProductField; "Hs" and "m0" the columns' significant
    wave height and zeroth moment; a callable is called with the model."""
    if callable(output):
        return np.asarray(output(model))
    if output in ("N", "action"):
        return model.action
    if output == "Hs":
        from .diagnostics import significant_wave_height
        return significant_wave_height(model)
    if output == "m0":
        from .diagnostics import m0
        return m0(model)
    raise ValueError(f"unknown output {output!r}: use 'N', 'Hs', 'm0' or a callable of the model")


class JLD2Writer:
    """`JLD2Writer(model, outputs; filename, schedule, overwrite_existing)`:
    outputs maps names to what fetch_output understands.  Register it in
    `simulation.output_writers`; `run_` writes at every actuation of the
    schedule, the starting iteration included."""

    def __init__(self, model, outputs, *, filename, schedule, overwrite_existing=False):
        if not filename.endswith(".jld2"):
            filename = filename + ".jld2"
        self.model, self.outputs, self.filename, self.schedule = model, dict(outputs), filename, schedule
        self.overwrite_existing = overwrite_existing
        self._file = None
        self._written = set()

    def _open(self):
        if self._file is None:
            self._file = JLD2File(self.filename, overwrite_existing=self.overwrite_existing)
            g = self.model.grid
            self._file["grid/Nx"] = np.int64(g.Nx)
            self._file["grid/Ny"] = np.int64(g.Ny)
            self._file["grid/Nz"] = np.int64(g.Nz)
            self._file["grid/xF"] = np.asarray(g.x_faces, dtype=np.float64)
            self._file["grid/yF"] = np.asarray(g.y_faces, dtype=np.float64)
            sg = self.model.spectral_grid
            self._file["spectral_grid/kappa"] = np.asarray(sg.kappa, dtype=np.float64)
            self._file["spectral_grid/phi"] = np.asarray(sg.phi, dtype=np.float64)
        return self._file

    def write(self, model):
        it = model.clock.iteration
        if it in self._written:
            return
        f = self._open()
        for name, out in self.outputs.items():
            f[f"timeseries/{name}/{it}"] = fetch_output(out, model)
        f[f"timeseries/t/{it}"] = float(model.clock.time)
        f.flush()
        self._written.add(it)

    def close(self):
        if self._file is not None:
            self._file.close()
            self._file = None
