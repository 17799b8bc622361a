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
    # split: members from that many tiles on run their own stage kernel
    # (-1 none, 1 every member, 40 the larger ones here)
    for ty, blocks, graph, split in ((4, 0, False, -1), (2, 64, False, 40), (8, 2048, True, -1), (4, 8, True, 40),
                                     (4, 0, True, 1)):
        ms = [build(w) for w in wls]
        ens = ra.Ensemble(ms)
        ens.set_launch(ty, blocks)
        ens.set_split(split)
        ens.use_graph(graph)
        ens.step(dt, nsteps - 1)
        ens.step(dt, 1)
        ens.synchronize()
        launch = ens.launch()
        if split == 40:
            assert 0 < launch["own_members"] < len(ms), launch
        for k, (m, r) in enumerate(zip(ms, ref)):
            assert m.clock.iteration == nsteps
            assert np.array_equal(m.action, r), (ty, blocks, graph, split, k, rel_l2(m.action, r))
        say(f"ty={ty} blocks={blocks} graph={graph} split={split}: {launch} bitwise equal")
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
    for ty, blocks, graph, split in ((4, 0, False, -1), (8, 64, True, 40)):
        ms = [build(w, current=False) for w in wls]
        ens = ra.Ensemble(ms)
        ens.set_launch(ty, blocks)
        ens.set_split(split)
        ens.use_graph(graph)
        ens.step(dt, nsteps)
        ens.synchronize()
        for k, (m, r) in enumerate(zip(ms, ref)):
            assert np.array_equal(m.action, r), (ty, blocks, graph, split, k, rel_l2(m.action, r))
        say(f"ty={ty} blocks={blocks} graph={graph} split={split}: {ens.launch()} bitwise equal")
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
    assert best["stage_ms"] > 0 and 0 <= best["own_members"] <= len(ms)
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


def _attach_networks(models, bins, seed=0, audit_stride=3, refresh=0):
    """Per-member 1155-style networks (bins + 3 -> 512 x 3 -> bins, GELU),
    random weights, with a wind envelope that sends some members to the plain
    bundle and an audit on every audit_stride-th column."""
    from ripple_amd import surrogate as sg
    nets = []
    cfg = sg.SurrogateConfig(rate_unit=1e-7, wind_lo=0.0, wind_hi=15.0, audit_stride=audit_stride,
                             audit_bound=0.5, refresh=refresh)
    for k, (m, nb) in enumerate(zip(models, bins)):
        net = sg.SourceNetwork(nb + 3, nb, hidden=512, hidden_layers=3, batch_capacity=64, seed=seed + k)
        sg.attach_surrogate(m, net, cfg)
        nets.append(net)
    return nets


@pytest.mark.parametrize("refresh", [0, 1])
def test_ensemble_with_surrogate_equals_members_stepped_alone(refresh):
    """Config 5's defining mode: every member takes its sources from its own
    512 x 3 network (plain bundle outside the wind envelope and on failed
    audits).  The ensemble equals the members stepped alone with the same
    networks, bitwise, for two launch shapes, graph on and off, with the
    rates inferred every stage or once per step."""
    say(f"ensemble with surrogate vs single, refresh={refresh}")
    wls = members(5, True, seed_base=700)
    bins = [w.N0.shape[2] * w.N0.shape[3] for w in wls]
    dt, nsteps = 0.02, 3
    ref = []
    for k, w in enumerate(wls):
        m = build(w)
        nets = _attach_networks([m], [bins[k]], seed=k, refresh=refresh)
        ra.time_step_(m, dt, nsteps)
        ref.append(m.action)
        del nets
    for ty, blocks, graph, split in ((4, 0, False, -1), (2, 64, True, 40)):
        ms = [build(w) for w in wls]
        nets = []
        for k, m in enumerate(ms):
            nets += _attach_networks([m], [bins[k]], seed=k, refresh=refresh)
        ens = ra.Ensemble(ms)
        ens.set_launch(ty, blocks)
        ens.set_split(split)
        ens.use_graph(graph)
        ens.step(dt, nsteps)
        ens.synchronize()
        for k, (m, r) in enumerate(zip(ms, ref)):
            assert np.array_equal(m.action, r), (ty, blocks, graph, split, k, rel_l2(m.action, r))
        say(f"surrogate ty={ty} blocks={blocks} graph={graph} split={split}: {ens.launch()} bitwise equal")
        ens.close()


def test_sampled_members_with_sources_against_fp64_oracle(oracle):
    """Four sampled config-5 members with the plain source bundle, stepped in
    the ensemble in FP32, against FP64 oracle runs of the same members."""
    from ripple_amd.sources import to_struct
    say("ensemble with sources vs fp64 oracle")
    wls = members(4, True, size_range=(12, 40), seed_base=800)
    ms = [build(w) for w in wls]
    ens = ra.Ensemble(ms)
    ens.use_graph(True)
    ens.step(0.02, 5)
    ens.close()
    for w, m in zip(wls, ms):
        src = oracle.sources_from(to_struct(w.model_kwargs["sources"]))
        om = oracle.OracleSourceModel(w, src, transport=oracle.MODE_REFRACTION)
        om.step(0.02, 5)
        err = rel_l2(m.action, om.action)
        say(f"{w.name}: FP32 ensemble with sources vs FP64 oracle rel-L2 {err:.2e}")
        assert err < 5e-3


def test_graph_recaptured_after_member_changes():
    """A member's sources change between graph-replayed ensemble steps, and a
    surrogate is attached to another: the ensemble recaptures and still
    equals the members stepped alone with the same changes."""
    from ripple_amd import sources as srcmod
    say("ensemble graph after member changes")
    wls = members(3, True, seed_base=900)
    bins = [w.N0.shape[2] * w.N0.shape[3] for w in wls]
    dt = 0.02
    calmer = ra.PrecomputedSources(
        wind_input=ra.PressureCorrelationInput(drag=ra.BulkWindDrag("linear"), wind=3.0, direction=1.0),
        dissipation=ra.LocalSaturationDissipation())

    def change(models, keep):
        srcmod.attach(models[0], calmer)
        keep += _attach_networks([models[1]], [bins[1]], seed=11)

    ref, keep = [build(w) for w in wls], []
    for m in ref:
        ra.time_step_(m, dt, 2)
    change(ref, keep)
    for m in ref:
        ra.time_step_(m, dt, 2)
    ms = [build(w) for w in wls]
    ens = ra.Ensemble(ms)
    ens.use_graph(True)
    ens.step(dt, 2)
    change(ms, keep)
    ens.step(dt, 2)
    ens.synchronize()
    for k, (m, r) in enumerate(zip(ms, ref)):
        assert np.array_equal(m.action, r.action), (k, rel_l2(m.action, r.action))
    ens.close()


def test_member_rules_rechecked_at_every_step():
    """A member that gains propagation smoothing after the ensemble was made
    is refused at the next ensemble step (the stage kernels cannot run it)."""
    say("ensemble member rules at step")
    wls = members(2, True, seed_base=950)
    ms = [build(w) for w in wls]
    ens = ra.Ensemble(ms)
    ens.step(0.02, 1)
    lib = ra._lib.load()
    ra._lib.check(lib.rpl_set_propagation_smoothing(ms[1]._ctx.ptr, 1, 0.5, 0.5))
    with pytest.raises(ra.RippleError):
        ens.step(0.02, 1)
    ra._lib.check(lib.rpl_set_propagation_smoothing(ms[1]._ctx.ptr, 0, 0.5, 0.5))
    ens.step(0.02, 1)
    ens.close()
