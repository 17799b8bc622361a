"""Benchmark of the wave-action time step on one MI355X.

A step is one SSP-RK3 time step (three fused tendency+stage kernel launches
plus the source preparation of each stage) of the translating-hurricane
configuration: 1024 x 1024 cells x 32 kappa x 36 phi bins, FP32 state, WENO5
intrinsic transport, the plain source bundle (Holland wind input with the
stress cap, local-saturation whitecapping); then the same step with the
learned 1155 -> 512 x 3 -> 1152 surrogate supplying the sources.  The action
is resident in HBM before timing starts.

Prints one JSON line: cell-bin updates per second per RK3 step with the plain
sources (the value) and with the surrogate, the roofline of the dominant
kernel (the stage kernel) measured with HIP events on the library's stream,
the other configurations' lines, the FP32 error against the FP64 path, and
the CPU oracle timed on a bounded sample of the same workload on this host.
"""
import argparse
import dataclasses
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ripple.jl_amd"))

import numpy as np  # noqa: E402

import ripple_amd as ra  # noqa: E402
from ripple_amd import workloads, _lib  # noqa: E402

METRIC = "cell-bin updates/s per RK3 step (plain vs surrogate), HBM GB/s; err vs FP64"
HBM_PEAK_GBS = 8000.0


def build_model(wl, float_type, kernel="auto", **extra):
    from ripple_amd import grids
    g, sg = wl.grid, wl.spectral_grid
    grid = grids.RectilinearGrid(float_type=float_type, size=(g.Nx, g.Ny, g.Nz), x=g.x_faces, y=g.y_faces,
                                 z=g.z_faces)
    if isinstance(sg, grids.FrequencyDirectionGrid):
        sgrid = grids.FrequencyDirectionGrid(float_type, frequency=sg.frequency, phi=sg.phi,
                                             frequency_faces=sg.frequency_faces, phi_faces=sg.phi_faces)
    else:
        sgrid = grids.PolarWaveVectorGrid(float_type, kappa=sg.kappa, phi=sg.phi, kappa_faces=sg.kappa_faces,
                                          phi_faces=sg.phi_faces)
    kw = dict(wl.model_kwargs)
    kw.setdefault("velocities", wl.velocities)
    kw.update(extra)
    m = ra.SpectralWaveModel(grid, sgrid, kernel=kernel, **kw)
    ra.set_(m, N=wl.N0)
    return m


def stage_bytes_per_launch(nx, ny, nk, nphi, esize):
    """Algorithmic HBM bytes of one stage launch, averaged over the three
    stages: N_k read and N_{k+1} written every stage, N0 read in stages 2-3,
    and the two Doppler fields U_x, U_y (the gradients are formed in-kernel)."""
    cells = nx * ny * nk * nphi
    per_step = esize * (8 * cells + 3 * 2 * nx * ny * nk)
    return per_step / 3.0


def intrinsic_stage_bytes_per_launch(nx, ny, nk, nphi, esize, column_fields=6):
    """Algorithmic HBM bytes of one intrinsic-transport stage launch with the
    source bundle, averaged over the three stages: N_k read and N_{k+1}
    written every stage, N0 read in stages 2-3, and the per-column source
    inputs (U10, u*, cos and sin of the wind angle, log z1, stress factor)."""
    cells = nx * ny * nk * nphi
    return esize * (8 * cells / 3.0 + column_fields * nx * ny)


def measured_traffic(workload, kernel="stage_phi_kernel"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (scaled by the
    copy-kernel factors recorded there), or None if there is none."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    for r in rec if isinstance(rec, list) else [rec]:
        if r.get("workload") == workload:
            return r["traffic_bytes_per_launch"]
    return None


def time_model(model, dt, steps, warmup):
    lib, ctx = model._ctx.lib, model._ctx.ptr
    ra.time_step_(model, dt, warmup)
    model.synchronize()
    _lib.check(lib.rpl_enable_kernel_timing(ctx, 1))
    model.synchronize()
    t0 = time.perf_counter()
    _lib.check(lib.rpl_time_step(ctx, float(dt), int(steps)))
    model.synchronize()
    t1 = time.perf_counter()
    import ctypes
    ms, n = ctypes.c_double(), ctypes.c_int64()
    _lib.check(lib.rpl_get_kernel_timing(ctx, ctypes.byref(ms), ctypes.byref(n)))
    _lib.check(lib.rpl_enable_kernel_timing(ctx, 0))
    return (t1 - t0) / steps, ms.value / max(n.value, 1), n.value


MLP_FLOP_PER_COLUMN = 2 * (1155 * 512 + 512 * 512 * 2 + 512 * 1152)  # 1155 -> 512 x3 -> 1152


def phase_ms(model, getter):
    """Average device milliseconds of one timing phase since its last read."""
    import ctypes
    ms, n = ctypes.c_double(), ctypes.c_int64()
    _lib.check(getattr(model._ctx.lib, getter)(model._ctx.ptr, ctypes.byref(ms), ctypes.byref(n)))
    return ms.value / max(n.value, 1), n.value


def storm_plain_vs_surrogate(steps, warmup, nx=1024, ny=1024):
    """The translating-hurricane configuration (1024 x 1024 cells x 32 kappa
    x 36 phi, FP32, intrinsic WENO5 transport, Holland wind input with stress
    cap and local-saturation whitecapping): RK3 step time with the plain
    source bundle and with the 1155 -> 512 x 3 -> 1152 network supplying the
    sources (the trained network shipped with the package, storm_surrogate;
    its configuration's terms and training envelope, outside which a column
    keeps the plain bundle), and the device time of each step's kernels (HIP
    events on the library stream)."""
    from ripple_amd import surrogate as sg, storm_surrogate
    wl = workloads.translating_hurricane(nx=nx, ny=ny)
    nx, ny, nk, nphi = wl.N0.shape
    m = build_model(wl, np.float32)
    del wl.N0
    t_plain, l_plain, n_plain = time_model(m, wl.dt, steps, warmup)
    l_state, n_state = phase_ms(m, "rpl_get_source_timing")
    net, cfg, hi, _ = sg.load_network(storm_surrogate.NETWORK_FILE, batch_capacity=64)
    sg.attach_surrogate(m, net, cfg, bin_action_hi=hi)
    ra.time_step_(m, wl.dt, 1)
    m.synchronize()
    phase_ms(m, "rpl_get_surrogate_timing")
    t_surr, l_surr, _ = time_model(m, wl.dt, steps, 0)
    t_inf, _ = phase_ms(m, "rpl_get_surrogate_timing")
    st = sg.surrogate_stats(m)
    # the same network with its rates inferred once per step and held
    # through the step's stages (an approximation; its accuracy is in
    # storm_surrogate_accuracy)
    sg.attach_surrogate(m, net, dataclasses.replace(cfg, refresh=1), bin_action_hi=hi)
    ra.time_step_(m, wl.dt, 1)
    m.synchronize()
    phase_ms(m, "rpl_get_surrogate_timing")
    t_once, _, _ = time_model(m, wl.dt, steps, 0)
    t_inf_once, _ = phase_ms(m, "rpl_get_surrogate_timing")
    flops = MLP_FLOP_PER_COLUMN * nx * ny
    cells = nx * ny * nk * nphi
    sg.attach_surrogate(m, None, None)
    net.close()
    return {"workload": wl.name, "cells": cells, "dt": wl.dt,
            "plain": {"value": cells / t_plain, "ms_per_step": t_plain * 1e3, "stage_kernel_ms": l_plain,
                      "stage_launches_timed": n_plain, "source_state_ms": l_state,
                      "source_state_launches_timed": n_state,
                      "other_ms_per_step": t_plain * 1e3 - 3 * (l_plain + l_state)},
            "surrogate": {"value": cells / t_surr, "ms_per_step": t_surr * 1e3, "stage_kernel_ms": l_surr,
                          "inference_ms_per_stage": t_inf,
                          "mlp_TFLOPs": flops / (t_inf * 1e-3) / 1e12,
                          "mlp_frac_of_bf16_dense_peak": flops / (t_inf * 1e-3) / 2.5e15,
                          "fallback_columns": st["plain"], "columns": st["columns"],
                          "network_terms": "every source term" if cfg.terms == 0 else "wind input (plain dissipation)"},
            "surrogate_rates_once_per_step": {"value": cells / t_once, "ms_per_step": t_once * 1e3,
                                              "inference_ms_per_step": t_inf_once,
                                              "step_speedup_over_plain": t_plain / t_once},
            "step_speedup_surrogate_over_plain": t_plain / t_surr}


def storm_surrogate_accuracy(nx=1024, ny=1024, nsteps=None, network_file=None):
    """The translating hurricane (nx x ny x 32 kappa x 36 phi) run for its
    nsteps (60 x 60 s) four ways from the same JONSWAP seed: FP32 with the
    plain source bundle, FP32 with the trained 1155 -> 512 x 3 -> 1152 network
    supplying every source term (plain bundle only where a column leaves the
    training envelope), FP64 with the plain bundle (the FP64 path parity-pinned
    to the oracle), and FP32 without sources (the scale of the source terms'
    effect).  Reports the significant-wave-height field's relative L2 error
    and the domain-mean m0's relative error of each run against the plain
    FP32 and the FP64 runs."""
    from ripple_amd import storm_surrogate, surrogate as sg, diagnostics as dg
    wl = workloads.translating_hurricane(nx=nx, ny=ny)
    nsteps = wl.nsteps if nsteps is None else nsteps
    net, cfg, hi, meta = sg.load_network(network_file or storm_surrogate.NETWORK_FILE, batch_capacity=64)
    out, fields = {}, {}

    def run(label, float_type, surrogate=False, sources=True, refresh=0):
        kw = {} if sources else {"sources": None}
        m = build_model(wl, float_type, **kw)
        if surrogate:
            sg.attach_surrogate(m, net, dataclasses.replace(cfg, refresh=refresh), bin_action_hi=hi)
        t0 = time.perf_counter()
        ra.time_step_(m, wl.dt, nsteps)
        m.synchronize()
        t = time.perf_counter() - t0
        hs = np.asarray(dg.significant_wave_height(m), dtype=np.float64)
        fields[label] = (hs, float(np.mean(np.asarray(dg.m0(m), dtype=np.float64))))
        rec = {"wall_s": t}
        if surrogate:
            rec["stats"] = sg.surrogate_stats(m)
            sg.attach_surrogate(m, None, None)
        out[label] = rec
        del m

    run("plain_f32", np.float32)
    run("surrogate_f32", np.float32, surrogate=True)
    run("surrogate_f32_rates_once_per_step", np.float32, surrogate=True, refresh=1)
    run("plain_f64", np.float64)
    run("no_sources_f32", np.float32, sources=False)
    net.close()
    for label in ("surrogate_f32", "surrogate_f32_rates_once_per_step", "plain_f32", "no_sources_f32"):
        hs, m0 = fields[label]
        for ref in ("plain_f32", "plain_f64"):
            if ref == label:
                continue
            hr, m0r = fields[ref]
            out[label][f"Hs_rel_l2_vs_{ref}"] = float(np.linalg.norm(hs - hr) / np.linalg.norm(hr))
            out[label][f"m0_rel_vs_{ref}"] = float(abs(m0 - m0r) / m0r)
            out[label][f"m0_ratio_to_{ref}"] = float(m0 / m0r)
    return {"workload": wl.name, "steps": nsteps, "network": os.path.basename(network_file or
                                                                               storm_surrogate.NETWORK_FILE),
            "training": {k: meta[k] for k in ("pairs", "held_out_rate_rel_l2") if k in meta}, **out}


def refraction_line(steps, warmup, nx, ny, float_type=np.float32, kernel="auto"):
    """The sheared-eddy refraction configuration at nx x ny x 32 kappa x 36
    phi (fused WENO5 Doppler transport + spectral refraction, no sources):
    RK3 step time and the stage kernel's roofline."""
    wl = workloads.sheared_eddies(nx=nx, ny=ny)
    m = build_model(wl, float_type, kernel=kernel)
    esize = np.dtype(float_type).itemsize
    t, l, n = time_model(m, wl.dt, steps, warmup)
    nk, nphi = wl.N0.shape[2], wl.N0.shape[3]
    bpl = stage_bytes_per_launch(nx, ny, nk, nphi, esize)
    ach = bpl / (l * 1e-3) / 1e9
    return {"workload": wl.name, "dtype": np.dtype(float_type).name, "kernel": kernel, "value": wl.cell_bins / t,
            "ms_per_step": t * 1e3, "avg_launch_ms": l, "launches_timed": n, "bytes_per_launch": bpl,
            "achieved_GBs": ach, "frac_of_hbm_peak": ach / HBM_PEAK_GBS,
            "traffic": measured_traffic(wl.name) if float_type == np.float32 else None}


def pseudomomentum_line(steps, warmup):
    """The self-coupled pseudomomentum current at the headline shape (512 x 512
    cells x 32 kappa x 36 phi, FP32): a swell packet carrying its own current,
    Q grid of 32 levels over 10 m.  Each stage first contracts its input action
    into U and dU/dkappa (pm_moment_kernel: reads N once, writes four
    (nx, ny, nkappa) planes), then runs the refraction stage kernel."""
    import ctypes
    wl = workloads.pseudomomentum_packet(float_type=np.float32, nx=512, ny=512, nz=32, nkappa=32, nphi=36,
                                         dt=0.01)
    nx, ny, nk, nphi = wl.N0.shape
    m = build_model(wl, np.float32)
    t_step, l_stage, _ = time_model(m, wl.dt, steps, warmup)
    ms, n = ctypes.c_double(), ctypes.c_int64()
    _lib.check(m._ctx.lib.rpl_get_coupling_timing(m._ctx.ptr, ctypes.byref(ms), ctypes.byref(n)))
    l_pm = ms.value / max(n.value, 1)
    pm_bytes = 4 * (nx * ny * nk * nphi + 4 * nx * ny * nk)
    return {"workload": wl.name, "value": wl.cell_bins / t_step, "ms_per_step": t_step * 1e3,
            "stage_kernel_ms": l_stage, "coupling_update_ms": l_pm, "coupling_updates_timed": n.value,
            "coupling_bytes_per_update": pm_bytes,
            "coupling_achieved_GBs": pm_bytes / (l_pm * 1e-3) / 1e9 if l_pm > 0 else None}


def smoothing_line(steps, warmup):
    """SpatialAveraging (alpha_s = alpha_n = 0.5) after every step of the
    headline workload: the step time with it, and its one pass per step
    (spatial_averaging_kernel: N read, the averaged field written)."""
    import ctypes
    wl = workloads.sheared_eddies()
    m = build_model(wl, np.float32, propagation_smoothing=ra.SpatialAveraging(0.5, 0.5))
    t_step, l_stage, _ = time_model(m, wl.dt, steps, warmup)
    ms, n = ctypes.c_double(), ctypes.c_int64()
    _lib.check(m._ctx.lib.rpl_get_smoothing_timing(m._ctx.ptr, ctypes.byref(ms), ctypes.byref(n)))
    l_sm = ms.value / max(n.value, 1)
    sm_bytes = 2 * 4 * wl.cell_bins
    return {"workload": wl.name + "_spatial_averaging", "value": wl.cell_bins / t_step, "ms_per_step": t_step * 1e3,
            "stage_kernel_ms": l_stage, "smoothing_ms": l_sm, "smoothing_passes_timed": n.value,
            "smoothing_bytes_per_pass": sm_bytes,
            "smoothing_achieved_GBs": sm_bytes / (l_sm * 1e-3) / 1e9 if l_sm > 0 else None}


def dia_line(steps, warmup):
    """The wind sea with the full source bundle including the DIA nonlinear
    transfer (256 x 256 x 40 frequencies x 36 directions, FP32, intrinsic
    transport): the RK3 step with and without the nonlinear member, and the
    DIA source launch (dia_source_kernel: the columns' action and transfer in
    LDS, the lookups tabulated)."""
    import ctypes
    from ripple_amd import sources as S
    wl = workloads.wind_sea_with_dia()
    m = build_model(wl, np.float32)
    t_step, l_stage, _ = time_model(m, wl.dt, steps, warmup)
    ms, n = ctypes.c_double(), ctypes.c_int64()
    _lib.check(m._ctx.lib.rpl_get_nonlinear_timing(m._ctx.ptr, ctypes.byref(ms), ctypes.byref(n)))
    l_dia = ms.value / max(n.value, 1)
    del m
    src = wl.model_kwargs["sources"]
    plain = S.PrecomputedSources(wind_input=src.wind_input, dissipation=src.dissipation)
    m2 = build_model(wl, np.float32, sources=plain)
    t_plain, _, _ = time_model(m2, wl.dt, steps, warmup)
    # What a network of the hurricane network's width costs per stage on this
    # grid (1443 -> 512 x 3 -> 1440, untrained weights: cost only), against
    # the DIA launch.  The network's rate form r * sigma * N cannot carry the
    # DIA's transfer into empty bins, so it stands in for the bundle's wind
    # input and dissipation only; the timing is what replacing the DIA by an
    # inference of this size would cost.
    from ripple_amd import surrogate as sg
    nb = wl.N0.shape[2] * wl.N0.shape[3]
    net = sg.SourceNetwork(nb + 3, nb, hidden=512, hidden_layers=3, batch_capacity=64, seed=1)
    sg.attach_surrogate(m2, net, sg.SurrogateConfig(terms=0))
    ra.time_step_(m2, wl.dt, 1)
    m2.synchronize()
    phase_ms(m2, "rpl_get_surrogate_timing")
    t_net, _, _ = time_model(m2, wl.dt, steps, 0)
    t_inf, _ = phase_ms(m2, "rpl_get_surrogate_timing")
    sg.attach_surrogate(m2, None, None)
    net.close()
    cells = wl.cell_bins
    return {"workload": wl.name, "value": cells / t_step, "ms_per_step": t_step * 1e3, "stage_kernel_ms": l_stage,
            "dia_source_ms": l_dia, "dia_launches_timed": n.value,
            "without_dia_ms_per_step": t_plain * 1e3,
            "dia_donor_pairs_per_s": 2 * cells / (l_dia * 1e-3) if l_dia > 0 else None,
            "network_inference_ms_per_stage": t_inf,
            "without_dia_with_network_ms_per_step": t_net * 1e3,
            "network": f"{nb + 3} -> 512 x 3 -> {nb}, untrained (cost only)"}


def ensemble_line(count, steps, with_surrogate=True, **member_kw):
    """The batched ensemble: `count` members drawn as ensemble_member(0..count-1)
    (64^2-512^2 cells x 16 levels, 16-40 kappa x 24-48 phi bins, own eddy
    current and uniform wind, wind input + local-saturation sources), FP32.
    Times the members stepped one after another (rpl_time_step each) and the
    ensemble (large members' own stage kernels on branch streams beside one
    persistent work-list launch per stage for the small ones; autotuned split
    and shape; hipGraph replay),
    first with the plain sources, then with every member's sources from a
    512 x 3 network of its own bin count (random weights, open envelope: a
    timing run)."""
    from ripple_amd import surrogate as sg
    models, bins, cells = [], [], 0
    for k in range(count):
        wl = workloads.ensemble_member(k, **member_kw)
        models.append(build_model(wl, np.float32))
        cells += wl.cell_bins
        bins.append(wl.N0.shape[2] * wl.N0.shape[3])
        del wl
    dt = 0.02

    def run(label):
        for m in models:
            ra.time_step_(m, dt, 1)
        for m in models:
            m.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            for m in models:
                ra.time_step_(m, dt, 1)
        for m in models:
            m.synchronize()
        t_seq = (time.perf_counter() - t0) / steps
        ens = ra.Ensemble(models)
        tuned = ens.autotune(dt, reps=2)
        ens.use_graph(True)
        ens.step(dt, 1)
        ens.synchronize()
        t0 = time.perf_counter()
        ens.step(dt, steps)
        ens.synchronize()
        t_ens = (time.perf_counter() - t0) / steps
        launch = ens.launch()
        ens.close()
        return {"sources": label,
                "sequential": {"value": cells / t_seq, "ms_per_step": t_seq * 1e3},
                "ensemble": {"value": cells / t_ens, "ms_per_step": t_ens * 1e3, "graph": True,
                             "autotuned": tuned, "launch": launch},
                "speedup_ensemble_over_sequential": t_seq / t_ens}

    out = {"members": count, "cell_bins": cells, "dtype": "f32", "steps": steps, "plain": run("plain bundle")}
    if member_kw:
        out["member_ranges"] = {k: list(v) for k, v in member_kw.items()}
    if not with_surrogate:
        return out
    nets = []
    for k, (m, nb) in enumerate(zip(models, bins)):
        net = sg.SourceNetwork(nb + 3, nb, hidden=512, hidden_layers=3, batch_capacity=64, seed=k)
        sg.attach_surrogate(m, net, sg.SurrogateConfig(rate_unit=1e-7))
        nets.append(net)
    out["surrogate"] = run("network, 512 x 3 per member")
    for m in models:
        sg.attach_surrogate(m, None, None)
    for n in nets:
        n.close()
    return out


def err_vs_fp64():
    """FP32 against FP64 after the 400-step vortex-refraction run (config 1)."""
    wl = workloads.vortex_refraction()
    m64 = build_model(wl, np.float64)
    m32 = build_model(wl, np.float32)
    for m in (m64, m32):
        ra.run_(ra.Simulation(m, dt=wl.dt, stop_iteration=wl.nsteps))
    N64, N32 = m64.action, m32.action.astype(np.float64)

    def rl2(a, b):
        return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))

    out = {"workload": wl.name + "_400_steps", "N_rel_l2": rl2(N32, N64)}
    out["m0_rel_l2"] = rl2(ra.m0(m32), ra.m0(m64))
    out["Hs_rel_l2"] = rl2(ra.significant_wave_height(m32), ra.significant_wave_height(m64))
    out["kappa_rms_rel_l2"] = rl2(ra.root_mean_square_wavenumber(m32), ra.root_mean_square_wavenumber(m64))
    d = np.angle(np.exp(1j * (ra.mean_direction(m32) - ra.mean_direction(m64))))
    out["mean_dir_max_abs_rad"] = float(np.max(np.abs(d)))
    a64, a32 = ra.total_action(m64), ra.total_action(m32)
    out["total_action_rel_diff"] = abs(a32 - a64) / abs(a64)
    return out


def err_vs_fp64_sheared_eddies(nsteps=200):
    """FP32 against FP64 after the sheared-eddy run at its full 512 x 512 x
    32 x 36 size (200 steps of dt = 0.02): N, m0, kappa_rms and mean
    direction."""
    wl = workloads.sheared_eddies()
    out = {"workload": f"{wl.name}_{nsteps}_steps"}
    m64 = build_model(wl, np.float64)
    ra.time_step_(m64, wl.dt, nsteps)
    N64 = m64.action
    d64 = (ra.m0(m64), ra.root_mean_square_wavenumber(m64), ra.mean_direction(m64), ra.total_action(m64))
    del m64
    m32 = build_model(wl, np.float32)
    ra.time_step_(m32, wl.dt, nsteps)
    N32 = m32.action

    def rl2(a, b):
        a = np.asarray(a, np.float64)
        return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))

    out["N_rel_l2"] = rl2(N32, N64)
    del N32, N64
    out["m0_rel_l2"] = rl2(ra.m0(m32), d64[0])
    out["kappa_rms_rel_l2"] = rl2(ra.root_mean_square_wavenumber(m32), d64[1])
    d = np.angle(np.exp(1j * (ra.mean_direction(m32) - d64[2])))
    out["mean_dir_max_abs_rad"] = float(np.max(np.abs(d)))
    out["total_action_rel_diff"] = abs(ra.total_action(m32) - d64[3]) / abs(d64[3])
    return out


def cpu_threads():
    """Host threads this process may use: its CPU affinity, capped at the
    OMP_NUM_THREADS the host sets for one GPU's share (when it sets one)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline():
    """The CPU oracle (FP64, the reference's operation order, OpenMP over
    (j, m, n)) on bounded samples of the headline workload: the hurricane's
    spectral grid (32 kappa x 36 phi), physics and source bundle on a
    32 x 32-column box with one thread and on a 128 x 128-column box with
    every host thread available; two RK3 steps after one warm-up each."""
    import platform
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from ripple_amd.sources import to_struct
    runs = {}
    for label, threads, n in (("single_thread", 1, 32), ("all_threads", cpu_threads(), 128)):
        oracle.set_threads(threads)
        wl = workloads.translating_hurricane(nx=n, ny=n, float_type=np.float64)
        src = oracle.sources_from(to_struct(wl.model_kwargs["sources"]))
        om = oracle.OracleSourceModel(wl, src, depth=200.0)
        om.step(wl.dt, 1)  # warm-up
        steps = 2
        t0 = time.perf_counter()
        om.step(wl.dt, steps)
        t = (time.perf_counter() - t0) / steps
        runs[label] = {"threads": threads, "columns": f"{n}x{n}", "ms_per_step": t * 1e3,
                       "value": wl.cell_bins / t}
        del om
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    a = runs["all_threads"]
    return {"value": a["value"], "unit": "cell-bin updates/s", "cores": a["threads"], "kind": "port",
            "sample": (f"translating hurricane physics and 32x36 spectral grid on {a['columns']} columns, FP64 "
                       f"oracle with the source bundle, 2 RK3 steps after 1 warm-up, OpenMP {a['threads']} "
                       f"threads, {a['ms_per_step']:.0f} ms/step; Julia threads: n/a (reference not runnable)"),
            "single_thread": runs["single_thread"], "cpu_model": cpu, "host": platform.node(),
            "compiler": "gcc -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off"}


def write_performance_tsv(path, record):
    """The measured lines as Validation.jl's performance summary
    (case, operation, seconds, bytes, description): seconds per RK3 step,
    bytes = the line's algorithmic HBM bytes per step where it has them."""
    from ripple_amd.validation import PerformanceMetric, write_performance_summary
    rows = [PerformanceMetric(record["config"]["workload"], "rk3_step_plain", record["ms_per_step"] * 1e-3,
                              3 * record["roofline"]["bytes_per_launch"],
                              "FP32 hurricane RK3 step, plain sources (headline)"),
            PerformanceMetric(record["config"]["workload"], "rk3_step_surrogate",
                              record["surrogate"]["ms_per_step"] * 1e-3, 0.0,
                              "FP32 hurricane RK3 step, learned surrogate sources")]
    e = record.get("extra", {})
    for key, desc in (("refraction_1024x1024", "FP32 sheared-eddy refraction RK3 step"),
                      ("refraction_512x512", "FP32 sheared-eddy refraction RK3 step"),
                      ("refraction_512x512_fp64", "FP64 sheared-eddy refraction RK3 step")):
        if key in e:
            rows.append(PerformanceMetric(e[key]["workload"], "rk3_step_" + e[key]["dtype"],
                                          e[key]["ms_per_step"] * 1e-3, 3 * e[key]["bytes_per_launch"], desc))
    for key, op, desc in (("pseudomomentum_coupling", "rk3_step", "RK3 step with the self-coupled pseudomomentum current"),
                          ("propagation_smoothing", "rk3_step", "RK3 step plus SpatialAveraging"),
                          ("nonlinear_dia", "rk3_step", "RK3 step with wind input, dissipation and DIA")):
        if key in e:
            rows.append(PerformanceMetric(e[key]["workload"], op, e[key]["ms_per_step"] * 1e-3, 0.0, desc))
    write_performance_summary(path, rows)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the FP64 and larger-grid lines")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble line")
    ap.add_argument("--only-pseudomomentum", action="store_true",
                    help="print only the self-coupled pseudomomentum line (for profiling its kernels)")
    ap.add_argument("--only-dia", action="store_true",
                    help="print only the DIA nonlinear-source line (for profiling its kernel)")
    ap.add_argument("--only-smoothing", action="store_true",
                    help="print only the propagation-smoothing line (for profiling its kernel)")
    ap.add_argument("--only-headline", action="store_true",
                    help="print only the plain vs surrogate hurricane line (for profiling its kernels)")
    ap.add_argument("--only-refraction", type=int, default=0, metavar="N",
                    help="print only the N x N sheared-eddy refraction line (for profiling its kernel)")
    ap.add_argument("--fp64-fast", action="store_true", help="with --only-refraction: FP64 on the fast kernel")
    ap.add_argument("--ensemble-members", type=int, default=64)
    ap.add_argument("--only-surrogate-accuracy", action="store_true",
                    help="print only the trained surrogate's accuracy on the 1024^2 hurricane")
    ap.add_argument("--only-ensemble", action="store_true",
                    help="print only the batched-ensemble lines (members stepped one by one vs batched)")
    ap.add_argument("--performance-summary", default=None,
                    help="also write the measured lines as a reference-format performance summary (TSV)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (reference for HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the wave step is built and measured on a single MI355X; "
              "multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4, "launches": 3}))
        return

    if args.only_pseudomomentum:
        print(json.dumps({"pseudomomentum_coupling": pseudomomentum_line(args.steps, args.warmup)}), flush=True)
        return
    if args.only_ensemble:
        print(json.dumps({"ensemble": ensemble_line(args.ensemble_members, max(2, args.steps // 5))}), flush=True)
        print(json.dumps({"ensemble_small_members": ensemble_line(256, 10, with_surrogate=False, size_range=(16, 64),
                                                                  kappa_range=(8, 16), phi_range=(12, 24))}),
              flush=True)
        return
    if args.only_dia:
        print(json.dumps({"nonlinear_dia": dia_line(args.steps, args.warmup)}), flush=True)
        return
    if args.only_smoothing:
        print(json.dumps({"propagation_smoothing": smoothing_line(args.steps, args.warmup)}), flush=True)
        return
    if args.only_surrogate_accuracy:
        print(json.dumps({"storm_surrogate_accuracy": storm_surrogate_accuracy()}), flush=True)
        return
    if args.only_headline:
        print(json.dumps({"storm_plain_vs_surrogate": storm_plain_vs_surrogate(args.steps, args.warmup)}), flush=True)
        return
    if args.only_refraction:
        n = args.only_refraction
        ft, kern = (np.float64, "fast") if args.fp64_fast else (np.float32, "auto")
        print(json.dumps({"refraction": refraction_line(args.steps, args.warmup, n, n, ft, kern)}), flush=True)
        return

    st = storm_plain_vs_surrogate(args.steps, args.warmup)
    nx, ny = 1024, 1024
    nk, nphi = 32, 36
    plain, surr = st["plain"], st["surrogate"]
    bpl = intrinsic_stage_bytes_per_launch(nx, ny, nk, nphi, 4)
    achieved = bpl / (plain["stage_kernel_ms"] * 1e-3) / 1e9
    record = {
        "metric": METRIC,
        "value": plain["value"],
        "unit": "cell-bin updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": plain["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": st["workload"], "nx": nx, "ny": ny, "nkappa": nk, "nphi": nphi, "nz": 32,
                   "dt": st["dt"], "transport": "fused WENO5 intrinsic transport",
                   "sources": "PrecomputedSources(PressureCorrelationInput(BulkWindDrag(:linear), Holland "
                              "hurricane on a LinearStormTrack) with stress cap, LocalSaturationDissipation)",
                   "timestepper": "SSP-RK3", "kernel": "fast"},
        "surrogate": {"value": surr["value"], "ms_per_step": surr["ms_per_step"],
                      "speedup_over_plain": st["step_speedup_surrogate_over_plain"],
                      "network": "1155 -> 512 x 3 -> 1152 GELU, bf16 weights and activations, fp32 accumulation",
                      "inference_ms_per_stage": surr["inference_ms_per_stage"], "mlp_TFLOPs": surr["mlp_TFLOPs"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(st["workload"]),
                     "kernel": "stage_phi_kernel (intrinsic transport + source bundle + RK3 stage)",
                     "bytes_per_launch": bpl, "avg_launch_ms": plain["stage_kernel_ms"],
                     "launches_timed": plain["stage_launches_timed"]},
        "kernels_per_stage_ms": {"stage_phi_kernel": plain["stage_kernel_ms"],
                                 "source_state_kernel": plain["source_state_ms"]},
    }
    extra = {"storm_plain_vs_surrogate": st}
    if not args.no_extra:
        extra["refraction_1024x1024"] = refraction_line(max(3, args.steps // 2), 1, 1024, 1024)
        extra["refraction_512x512"] = refraction_line(args.steps, args.warmup, 512, 512)
        extra["refraction_512x512_fp64"] = refraction_line(max(3, args.steps // 4), 1, 512, 512, np.float64)
        extra["refraction_512x512_fp64_fast"] = refraction_line(max(3, args.steps // 2), 1, 512, 512, np.float64,
                                                                kernel="fast")
        extra["err_vs_fp64"] = err_vs_fp64()
        extra["err_vs_fp64_sheared_eddies"] = err_vs_fp64_sheared_eddies()
        extra["storm_surrogate_accuracy"] = storm_surrogate_accuracy()
        extra["pseudomomentum_coupling"] = pseudomomentum_line(max(3, args.steps // 2), 1)
        extra["propagation_smoothing"] = smoothing_line(max(3, args.steps // 2), 1)
        extra["nonlinear_dia"] = dia_line(max(3, args.steps // 4), 1)
    if not args.no_ensemble:
        extra["ensemble"] = ensemble_line(args.ensemble_members, max(2, args.steps // 5))
        extra["ensemble_small_members"] = ensemble_line(256, 10, with_surrogate=False, size_range=(16, 64),
                                                        kappa_range=(8, 16), phi_range=(12, 24))
    record["extra"] = extra
    if args.performance_summary:
        write_performance_tsv(args.performance_summary, record)
    if not args.no_cpu:
        record["cpu_baseline"] = cpu_baseline()
    print(json.dumps(record), flush=True)


if __name__ == "__main__":
    main()
