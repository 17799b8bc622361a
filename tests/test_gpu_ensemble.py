"""Batched ensembles: members of uneven physical and spectral sizes advanced
by one persistent launch per RK3 stage must equal the same models stepped one
by one (same arithmetic per cell, so bitwise), for every launch shape, with
and without the hipGraph replay and with source terms; sampled members are
checked against the FP64 oracle."""
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
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def build(wl, float_type=np.float32, current=True):
    g, sg = wl.grid, wl.spectral_grid
    grid = grids.RectilinearGrid(float_type=float_type, size=(g.Nx, g.Ny, g.Nz), x=g.x_faces, y=g.y_faces,
                                 z=g.z_faces)
    sgrid = grids.PolarWaveVectorGrid(float_type, kappa=sg.kappa, phi=sg.phi, kappa_faces=sg.kappa_faces,
                                      phi_faces=sg.phi_faces)
    # without the current the model takes the intrinsic WENO transport
    m = ra.SpectralWaveModel(grid, sgrid, velocities=wl.velocities if current else None, **wl.model_kwargs)
    ra.set_(m, N=wl.N0)
    return m


def members(count, sources, size_range=(9, 70), kappa_range=(3, 9), phi_range=(5, 12), seed_base=100):
    return [workloads.ensemble_member(k, nz=4, sources=sources, size_range=size_range, kappa_range=kappa_range,
                                      phi_range=phi_range, seed_base=seed_base) for k in range(count)]


@pytest.mark.parametrize("sources", [False, True])
def test_ensemble_equals_members_stepped_alone(sources):
    say(f"ensemble vs single, sources={sources}")
    wls = members(6, sources)
    say("sizes " + ", ".join(w.name for w in wls))
    dt, nsteps = 0.02, 3
    ref = []
    for w in wls:
        m = build(w)
        ra.time_step_(m, dt, nsteps)
        ref.append(m.action)
    for ty, blocks, graph in ((4, 0, False), (2, 64, False), (8, 2048, True), (4, 8, True)):
        ms = [build(w) for w in wls]
        ens = ra.Ensemble(ms)
        ens.set_launch(ty, blocks)
        ens.use_graph(graph)
        ens.step(dt, nsteps - 1)
        ens.step(dt, 1)
        ens.synchronize()
        for k, (m, r) in enumerate(zip(ms, ref)):
            assert m.clock.iteration == nsteps
            assert np.array_equal(m.action, r), (ty, blocks, graph, k, rel_l2(m.action, r))
        say(f"ty={ty} blocks={blocks} graph={graph}: {ens.launch()} bitwise equal")
        ens.close()


def test_intrinsic_ensemble_equals_members_stepped_alone():
    """Members without a current (intrinsic WENO transport, the per-bin work
    list) with sources: bitwise equal to the members stepped alone."""
    say("intrinsic ensemble vs single")
    wls = members(5, True, seed_base=500)
    dt, nsteps = 0.02, 3
    ref = []
    for w in wls:
        m = build(w, current=False)
        assert m.transport == ra._lib.RPL_TRANSPORT_INTRINSIC
        ra.time_step_(m, dt, nsteps)
        ref.append(m.action)
    for ty, blocks, graph in ((4, 0, False), (8, 64, True)):
        ms = [build(w, current=False) for w in wls]
        ens = ra.Ensemble(ms)
        ens.set_launch(ty, blocks)
        ens.use_graph(graph)
        ens.step(dt, nsteps)
        ens.synchronize()
        for k, (m, r) in enumerate(zip(ms, ref)):
            assert np.array_equal(m.action, r), (ty, blocks, graph, k, rel_l2(m.action, r))
        say(f"ty={ty} blocks={blocks} graph={graph}: bitwise equal")
        ens.close()


def test_autotune_keeps_results_and_state():
    say("ensemble autotune")
    wls = members(4, True, seed_base=200)
    ms = [build(w) for w in wls]
    before = [m.action for m in ms]
    ens = ra.Ensemble(ms)
    best = ens.autotune(0.02, reps=2)
    say(f"autotune {best}")
    assert best["tile_rows"] in (2, 4, 8) and best["blocks"] % 8 == 0 and best["blocks"] >= 0
    assert best["stage_ms"] > 0
    for m, b in zip(ms, before):
        assert np.array_equal(m.action, b)  # tuning leaves the state alone
    ens.use_graph(True)
    ens.step(0.02, 2)
    ens.close()
    for w, m in zip(wls, ms):
        r = build(w)
        ra.time_step_(r, 0.02, 2)
        assert np.array_equal(m.action, r.action)


def test_sampled_members_against_fp64_oracle(oracle):
    """FP32 ensemble members against FP64 oracle runs of the same members
    (refraction transport, no sources)."""
    say("ensemble vs fp64 oracle")
    wls = members(4, False, size_range=(12, 40), seed_base=300)
    ms = [build(w) for w in wls]
    ens = ra.Ensemble(ms)
    ens.step(0.02, 5)
    ens.close()
    for w, m in zip(wls, ms):
        om = oracle.OracleModel(w)
        om.step(0.02, 5)
        err = rel_l2(m.action, om.action)
        say(f"{w.name}: FP32 ensemble vs FP64 oracle rel-L2 {err:.2e}")
        assert err < 5e-3


def test_ensemble_error_paths():
    say("ensemble errors")
    w32 = members(2, False, seed_base=400)
    a, b = build(w32[0]), build(w32[1])
    c64 = build(w32[0], np.float64)
    with pytest.raises(ra.RippleError):
        ra.Ensemble([a, c64])
    with pytest.raises(ra.RippleArgumentError):
        ra.Ensemble([a, a])
    ws = members(1, True, seed_base=400)
    with pytest.raises(ra.RippleError):
        ra.Ensemble([a, build(ws[0])])  # sources on one member only
    ens = ra.Ensemble([a, b])
    with pytest.raises(ra.RippleArgumentError):
        ens.set_launch(3, 0)
    with pytest.raises(ra.RippleArgumentError):
        ens.set_launch(4, 12)
    with pytest.raises(ra.RippleArgumentError):
        ens.step(-1.0)
    ens.close()
