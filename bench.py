"""Benchmark of the wave-action time step on one MI355X.

A step is one SSP-RK3 time step (three fused tendency+stage kernel launches)
of the sheared-eddy configuration: 512 x 512 cells x 32 kappa x 36 phi bins,
FP32 state, WENO5 transport with Doppler shift and spectral refraction,
synthetic eddy current.  The action is resident in HBM before timing starts.

Prints one JSON line: cell-bin updates per second per RK3 step (the value),
the roofline of the stage kernel measured with HIP events on the library's
stream, the FP32 error against the FP64 path, and the CPU oracle timed on a
bounded sample of the same workload on this host.
"""
import argparse
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


def measured_traffic(workload):
    """HBM bytes per stage launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (scaled by the copy-kernel factors
    recorded there), or None if there is none for this workload."""
    path = os.path.join(ROOT, "profiles", "stage_phi_kernel_hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("workload") != workload:
        return None
    return rec["traffic_bytes_per_launch"]


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


def storm_plain_vs_surrogate(steps, warmup):
    """The translating-hurricane configuration at full size (1024 x 1024 cells
    x 32 kappa x 36 phi, FP32, intrinsic WENO5 transport, Holland wind input
    with stress cap and local-saturation whitecapping): RK3 step time with the
    plain source bundle and with the 1155 -> 512 x 3 -> 1152 network supplying
    the sources on every column.  The network weights are random (a timing
    run: the work per step does not depend on them); the envelope is open so
    that no column falls back."""
    import ctypes
    from ripple_amd import surrogate as sg
    wl = workloads.translating_hurricane()
    nx, ny, nk, nphi = wl.N0.shape
    nb = nk * nphi
    m = build_model(wl, np.float32)
    del wl.N0
    t_plain, l_plain, _ = time_model(m, wl.dt, steps, warmup)
    net = sg.SourceNetwork(nb + 3, nb, hidden=512, hidden_layers=3, batch_capacity=64, seed=1)
    cfg = sg.SurrogateConfig(rate_unit=1e-7)
    sg.attach_surrogate(m, net, cfg)
    lib, ctx = m._ctx.lib, m._ctx.ptr
    ra.time_step_(m, wl.dt, 1)
    m.synchronize()
    _lib.check(lib.rpl_get_surrogate_timing(ctx, None, None))
    t_surr, l_surr, _ = time_model(m, wl.dt, steps, 0)
    ms, n = ctypes.c_double(), ctypes.c_int64()
    _lib.check(lib.rpl_get_surrogate_timing(ctx, ctypes.byref(ms), ctypes.byref(n)))
    st = sg.surrogate_stats(m)
    t_inf = ms.value / max(n.value, 1)
    flops = MLP_FLOP_PER_COLUMN * nx * ny
    cells = nx * ny * nk * nphi
    sg.attach_surrogate(m, None, None)
    net.close()
    return {"workload": f"translating_hurricane_{nx}x{ny}_k{nk}_p{nphi}", "dtype": "f32 state, bf16 network",
            "plain": {"value": cells / t_plain, "ms_per_step": t_plain * 1e3, "stage_kernel_ms": l_plain},
            "surrogate": {"value": cells / t_surr, "ms_per_step": t_surr * 1e3, "stage_kernel_ms": l_surr,
                          "inference_ms_per_stage": t_inf,
                          "mlp_TFLOPs": flops / (t_inf * 1e-3) / 1e12,
                          "mlp_frac_of_bf16_dense_peak": flops / (t_inf * 1e-3) / 2.5e15,
                          "fallback_columns": st["plain"], "columns": st["columns"]},
            "step_speedup_surrogate_over_plain": t_plain / t_surr}


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
    cells = wl.cell_bins
    return {"workload": wl.name, "value": cells / t_step, "ms_per_step": t_step * 1e3, "stage_kernel_ms": l_stage,
            "dia_source_ms": l_dia, "dia_launches_timed": n.value,
            "without_dia_ms_per_step": t_plain * 1e3,
            "dia_donor_pairs_per_s": 2 * cells / (l_dia * 1e-3) if l_dia > 0 else None}


def ensemble_line(count, steps, with_surrogate=True, **member_kw):
    """The batched ensemble: `count` members drawn as ensemble_member(0..count-1)
    (64^2-512^2 cells x 16 levels, 16-40 kappa x 24-48 phi bins, own eddy
    current and uniform wind, wind input + local-saturation sources), FP32.
    Times the members stepped one after another (rpl_time_step each) and the
    ensemble (persistent launch per stage, autotuned shape, hipGraph replay),
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


def cpu_baseline(threads):
    """The CPU oracle (FP32, reference operation order, OpenMP) on a bounded
    sample: the same spectral grid, current and physics on a 96 x 96 sub-box."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.set_threads(threads)
    wl = workloads.sheared_eddies(nx=96, ny=96)
    om = oracle.OracleModel(wl, dtype=np.float32)
    om.step(wl.dt, 1)  # warm-up
    steps = 2
    t0 = time.perf_counter()
    om.step(wl.dt, steps)
    t = (time.perf_counter() - t0) / steps
    return {"value": wl.cell_bins / t, "unit": "cell-bin updates/s", "cores": threads, "kind": "port",
            "sample": f"{wl.name} (96x96 sub-box of the 512x512 workload), FP32, {steps} RK3 steps after 1 "
                      f"warm-up, OpenMP {threads} threads, {t * 1e3:.0f} ms/step"}


def write_performance_tsv(path, record):
    """The measured lines as Validation.jl's performance summary
    (case, operation, seconds, bytes, description): seconds per RK3 step,
    bytes = the line's algorithmic HBM bytes per step where it has them."""
    from ripple_amd.validation import PerformanceMetric, write_performance_summary
    rows = [PerformanceMetric(record["config"]["workload"], "rk3_step", record["ms_per_step"] * 1e-3,
                              3 * record["roofline"]["bytes_per_launch"], "FP32 sheared-eddy RK3 step (headline)")]
    e = record.get("extra", {})
    if "fp64_reference_kernel" in e:
        rows.append(PerformanceMetric(record["config"]["workload"], "rk3_step_fp64",
                                      e["fp64_reference_kernel"]["ms_per_step"] * 1e-3, 0.0,
                                      "FP64 reference-order RK3 step"))
    for key, op, desc in (("pseudomomentum_coupling", "rk3_step", "RK3 step with the self-coupled pseudomomentum current"),
                          ("propagation_smoothing", "rk3_step", "RK3 step plus SpatialAveraging"),
                          ("nonlinear_dia", "rk3_step", "RK3 step with wind input, dissipation and DIA")):
        if key in e:
            rows.append(PerformanceMetric(e[key]["workload"], op, e[key]["ms_per_step"] * 1e-3, 0.0, desc))
    if "storm_plain_vs_surrogate" in e:
        st = e["storm_plain_vs_surrogate"]
        rows.append(PerformanceMetric(st["workload"], "rk3_step_plain", st["plain"]["ms_per_step"] * 1e-3, 0.0,
                                      "hurricane RK3 step, plain sources"))
        rows.append(PerformanceMetric(st["workload"], "rk3_step_surrogate", st["surrogate"]["ms_per_step"] * 1e-3,
                                      0.0, "hurricane RK3 step, learned surrogate sources"))
    write_performance_summary(path, rows)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the FP64 and larger-grid lines")
    ap.add_argument("--no-storm", action="store_true", help="skip the plain vs surrogate hurricane line")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble line")
    ap.add_argument("--only-pseudomomentum", action="store_true",
                    help="print only the self-coupled pseudomomentum line (for profiling its kernels)")
    ap.add_argument("--only-dia", action="store_true",
                    help="print only the DIA nonlinear-source line (for profiling its kernel)")
    ap.add_argument("--only-smoothing", action="store_true",
                    help="print only the propagation-smoothing line (for profiling its kernel)")
    ap.add_argument("--ensemble-members", type=int, default=64)
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
    if args.only_dia:
        print(json.dumps({"nonlinear_dia": dia_line(args.steps, args.warmup)}), flush=True)
        return
    if args.only_smoothing:
        print(json.dumps({"propagation_smoothing": smoothing_line(args.steps, args.warmup)}), flush=True)
        return

    wl = workloads.sheared_eddies()
    nx, ny, nk, nphi = wl.N0.shape
    model = build_model(wl, np.float32)
    t_step, t_launch, launches = time_model(model, wl.dt, args.steps, args.warmup)
    value = wl.cell_bins / t_step
    bpl = stage_bytes_per_launch(nx, ny, nk, nphi, 4)
    achieved = bpl / (t_launch * 1e-3) / 1e9
    record = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-bin updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl.name, "nx": nx, "ny": ny, "nkappa": nk, "nphi": nphi, "nz": wl.grid.Nz,
                   "dt": wl.dt, "transport": "fused WENO5 Doppler transport + refraction",
                   "sources": None, "timestepper": "SSP-RK3", "kernel": "fast"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(wl.name),
                     "kernel": "stage_phi_kernel", "bytes_per_launch": bpl,
                     "avg_launch_ms": t_launch, "launches_timed": launches},
    }
    extra = {}
    if not args.no_extra:
        del model
        m64 = build_model(wl, np.float64)
        t64, l64, _ = time_model(m64, wl.dt, max(3, args.steps // 4), 1)
        extra["fp64_reference_kernel"] = {
            "value": wl.cell_bins / t64, "ms_per_step": t64 * 1e3, "avg_launch_ms": l64,
            "achieved_GBs": stage_bytes_per_launch(nx, ny, nk, nphi, 8) / (l64 * 1e-3) / 1e9}
        del m64
        big = workloads.sheared_eddies(nx=1024, ny=1024)
        mb = build_model(big, np.float32)
        tb, lb, _ = time_model(mb, big.dt, max(3, args.steps // 2), 1)
        extra["fp32_1024x1024"] = {
            "workload": big.name, "value": big.cell_bins / tb, "ms_per_step": tb * 1e3, "avg_launch_ms": lb,
            "achieved_GBs": stage_bytes_per_launch(1024, 1024, nk, nphi, 4) / (lb * 1e-3) / 1e9}
        del mb
        extra["err_vs_fp64"] = err_vs_fp64()
        extra["pseudomomentum_coupling"] = pseudomomentum_line(max(3, args.steps // 2), 1)
        extra["propagation_smoothing"] = smoothing_line(max(3, args.steps // 2), 1)
        extra["nonlinear_dia"] = dia_line(max(3, args.steps // 4), 1)
    if not args.no_storm:
        extra["storm_plain_vs_surrogate"] = storm_plain_vs_surrogate(max(2, args.steps // 4), 1)
    if not args.no_ensemble:
        extra["ensemble"] = ensemble_line(args.ensemble_members, max(2, args.steps // 5))
        extra["ensemble_small_members"] = ensemble_line(256, 10, with_surrogate=False, size_range=(16, 64),
                                                        kappa_range=(8, 16), phi_range=(12, 24))
    record["extra"] = extra
    if args.performance_summary:
        write_performance_tsv(args.performance_summary, record)
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        record["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(record), flush=True)


if __name__ == "__main__":
    main()
