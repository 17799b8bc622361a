"""JLD2 output of a Simulation (the reference's JLD2Writer use,
test/integration/model_api.jl:49-60, with fetch_output of the wave action
as its interior, src/OceananigansIntegration.jl:7-10), read back by the
HDF5 project's h5ls/h5dump."""
import math
import sys

import numpy as np
import pytest

import ripple_amd as ra
from ripple_amd import grids, workloads
import h5dump_reader as h5

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not h5.available(), reason="h5ls/h5dump not installed")]


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def test_reference_writer_case(tmp_path):
    """model_api.jl:49-60: one iteration with IterationInterval(1) writes the
    initial and the stepped action."""
    say("JLD2Writer, one iteration, no transport")
    grid = grids.RectilinearGrid(size=(4, 3, 2), x=(0, 4), y=(0, 3), z=(-1, 0))
    sg = grids.PolarWaveVectorGrid(kappa=np.array([0.4, 0.6, 0.8]), phi=np.arange(4) * (math.pi / 2))
    model = ra.SpectralWaveModel(grid, sg, horizontal_advection=None)
    ra.set_(model, N=1.0)
    path = str(tmp_path / "out.jld2")
    sim = ra.Simulation(model, dt=0.01, stop_iteration=1)
    sim.output_writers["fields"] = ra.JLD2Writer(model, {"N": "N"}, filename=path,
                                                 schedule=ra.IterationInterval(1), overwrite_existing=True)
    ra.run_(sim)
    assert model.clock.iteration == 1
    ls = h5.listing(path)
    assert ls["/timeseries/N/0"] == (4, 3, 3, 4) and ls["/timeseries/N/1"] == (4, 3, 3, 4)
    assert np.array_equal(h5.julia_array(path, "/timeseries/N/1"), model.action)
    assert np.array_equal(h5.julia_array(path, "/timeseries/N/0"), np.ones(model.size))
    assert h5.julia_array(path, "/timeseries/t/0") == 0.0
    assert h5.julia_array(path, "/timeseries/t/1") == pytest.approx(0.01)
    assert h5.julia_array(path, "/grid/Nx") == 4


def test_schedules_and_wave_height(tmp_path):
    """The vortex refraction run with N every 2 iterations and Hs on a time
    interval of 0.15 (three steps of 0.05); every snapshot equals the model's
    state at that iteration."""
    say("JLD2Writer, vortex refraction, IterationInterval(2) and TimeInterval(0.15)")
    wl = workloads.vortex_refraction(nsteps=6)
    model = ra.SpectralWaveModel(wl.grid, wl.spectral_grid, velocities=wl.velocities, **wl.model_kwargs)
    ra.set_(model, N=wl.N0)
    path_n, path_h = str(tmp_path / "n.jld2"), str(tmp_path / "hs.jld2")
    sim = ra.Simulation(model, dt=wl.dt, stop_iteration=6)
    sim.output_writers["N"] = ra.JLD2Writer(model, {"N": "N"}, filename=path_n, schedule=ra.IterationInterval(2))
    sim.output_writers["Hs"] = ra.JLD2Writer(model, {"Hs": "Hs"}, filename=path_h,
                                             schedule=ra.TimeInterval(3 * wl.dt))
    ra.run_(sim)
    assert model.clock.iteration == 6
    ls_n, ls_h = h5.listing(path_n), h5.listing(path_h)
    assert sorted(k for k in ls_n if k.startswith("/timeseries/N/")) == [f"/timeseries/N/{i}" for i in (0, 2, 4, 6)]
    assert sorted(k for k in ls_h if k.startswith("/timeseries/Hs/")) == [f"/timeseries/Hs/{i}" for i in (0, 3, 6)]
    # the last snapshots are the model's final state
    assert np.array_equal(h5.julia_array(path_n, "/timeseries/N/6"), model.action)
    np.testing.assert_array_equal(h5.julia_array(path_h, "/timeseries/Hs/6"), ra.significant_wave_height(model))
    # the initial snapshot is the initial condition
    np.testing.assert_array_equal(h5.julia_array(path_n, "/timeseries/N/0"), wl.N0)
    # a replay to iteration 2 gives the stored snapshot bit for bit
    replay = ra.SpectralWaveModel(wl.grid, wl.spectral_grid, velocities=wl.velocities, **wl.model_kwargs)
    ra.set_(replay, N=wl.N0)
    ra.time_step_(replay, wl.dt, 2)
    assert np.array_equal(h5.julia_array(path_n, "/timeseries/N/2"), replay.action)
    assert h5.julia_array(path_n, "/timeseries/t/4") == pytest.approx(4 * wl.dt)
